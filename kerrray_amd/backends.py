"""Backend protocol, registry and the ``hip`` backend.

The protocol and the registry keep the reference's contract
(src/kerrray/raytracing/backends/base.py:40-111): ``Backend`` has a
``name`` and ``integrate_batch(st, Y0, integ, term, events=None) ->
BatchResult``; ``get_backend(name)`` raises ``ValueError`` for an unknown
name and ``BackendUnavailable`` (a ``RuntimeError``) for a known one that
cannot run here. ``"hip"`` is added to the reference's names.

:class:`HipBackend` is the MI355X backend: the host packs the kernel
constants as the reference's Numba wrapper does
(numba_backend.py:87-128) and hands them, with the float64 initial states,
to ``kr_integrate_batch`` in ``libkrhip.so``, where one GPU lane integrates
one photon at a time (Dormand-Prince batches of 2^15 rows and more, 2^17 in
float32 or with the disk event, on persistent kernels that hand out the
longest photons first; the rows are the same either way). This package
implements only ``hip``; the reference's ``numpy``
and ``numba`` backends belong to the reference package and are reported
unavailable here rather than emulated.
"""

from __future__ import annotations

import ctypes
import math
import time
from typing import Final, Protocol, runtime_checkable

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd import _lib
from kerrray_amd.geodesics import (
    DP_A, DP_E, ERROR_EXPONENT, FACTOR_MAX, FACTOR_MIN, RK4_B, RK4_C, SAFETY, STATE_SIZE,
    EventOptions, IntegratorOptions, TerminationOptions, max_attempts,
)
from kerrray_amd.geometry import Spacetime, outer_horizon
from kerrray_amd.photons import AXIS_SIN_TOLERANCE, THETA_DOMAIN_TOLERANCE, BatchResult, ResultRows

__all__ = [
    "BACKEND_NAMES", "Backend", "BackendUnavailable", "HipBackend", "available_backends", "get_backend",
    "kernel_params",
]

BACKEND_NAMES: Final[tuple[str, ...]] = ("numpy", "numba", "cuda", "hip")
SUPPORTED_BACKENDS: Final[frozenset[str]] = frozenset(BACKEND_NAMES)
CUDA_UNAVAILABLE_MESSAGE: Final[str] = "no CUDA device is used; the GPU backend of this package is 'hip'"
REFERENCE_ONLY_MESSAGE: Final[str] = (
    "backend {name!r} belongs to the reference package kerrray; this package provides only 'hip'"
)


class BackendUnavailable(RuntimeError):
    """A known backend that cannot be used in this environment."""


@runtime_checkable
class Backend(Protocol):
    name: str

    def integrate_batch(
        self,
        st: Spacetime,
        Y0: ArrayLike,
        integ: IntegratorOptions,
        term: TerminationOptions,
        events: EventOptions | None = None,
    ) -> BatchResult: ...


def kernel_params(
    st: Spacetime, integ: IntegratorOptions, term: TerminationOptions, events: EventOptions | None = None
) -> _lib.KrParams:
    """The ``kr_params`` block: ``_kernel_constants`` + ``_tableaus`` of numba_backend.py:87-128.

    Products are written in the reference's order (``-2.0 * m * a`` is
    ``(-2.0 * m) * a``; ``a**2`` is Python's float power) so that the kernel
    sees the same bits the reference's kernel sees.
    """
    m, a = st.mass, st.a
    p = _lib.KrParams()
    geom = [m, a, a**2, -2.0 * m * a, 2.0 * m * a, 2.0 * m, 2.0 * a**2]
    r_in = events.r_in if events is not None else 0.0
    r_out = events.r_out if events is not None else 0.0
    consts = [
        0.0, 1.0, 2.0, -0.5, 8.0, SAFETY, FACTOR_MIN, FACTOR_MAX, ERROR_EXPONENT,
        integ.rtol, integ.atol, integ.h_min, outer_horizon(st) + term.horizon_epsilon,
        term.escape_radius, -THETA_DOMAIN_TOLERANCE, math.pi + THETA_DOMAIN_TOLERANCE,
        AXIS_SIN_TOLERANCE, integ.step_size, r_in, r_out, 0.5 * math.pi,
    ]
    h_max = math.inf if integ.h_max is None else integ.h_max
    params = [integ.lambda_max, 0.5 * math.pi, h_max]
    disk = events is not None and events.disk_plane
    ipar = [integ.max_steps, max_attempts(integ), int(integ.method == "rk45"), int(disk)]
    for i, v in enumerate(geom):
        p.geom[i] = float(v)
    for i, v in enumerate(consts):
        p.consts[i] = float(v)
    for i, v in enumerate(params):
        p.params[i] = float(v)
    for i, v in enumerate(ipar):
        p.ipar[i] = int(v)
    for i, row in enumerate(DP_A):
        for j, v in enumerate(row):
            p.tab_a[i * 7 + j] = v
    for i, v in enumerate(DP_E):
        p.tab_e[i] = v
    for i in range(4):
        p.rk4_b[i] = RK4_B[i]
        p.rk4_c[i] = RK4_C[i]
    p.dtype = _lib.KR_FLOAT64 if integ.dtype == "float64" else _lib.KR_FLOAT32
    return p


def _validate_batch(Y0: ArrayLike) -> NDArray[np.float64]:
    arr = np.array(Y0, dtype=np.float64)
    if arr.ndim != 2 or arr.shape[1] != STATE_SIZE:
        raise ValueError(f"Y0 must have shape (N, {STATE_SIZE}), got {arr.shape}")
    if not np.all(np.isfinite(arr)):
        raise ValueError("Y0 must be finite")
    return np.ascontiguousarray(arr)


def output_struct(rows: ResultRows) -> _lib.KrOutputs:
    """The ``kr_outputs`` block pointing at the (C-contiguous) arrays of ``rows``."""
    o = _lib.KrOutputs()
    for name, _ in _lib.KrOutputs._fields_:  # the struct's fields are the store's array names
        arr = getattr(rows, name)
        if arr is not None:
            if not arr.flags.c_contiguous:
                raise ValueError(f"output array {name} must be C-contiguous")
            setattr(o, name, arr.ctypes.data)
    return o


class HostBatch(ResultRows):
    """Freshly allocated host outputs of one batch and their ``kr_outputs`` view (``outputs``)."""

    def __init__(self, n: int, dtype: np.dtype, disk: bool) -> None:
        super().__init__(n, dtype, disk)
        self.disk = disk
        self.outputs = output_struct(self)


class HipBackend:
    """The MI355X backend (``name == "hip"``) behind the ``Backend`` protocol.

    Construction opens a handle on ``device``; a missing library or device
    raises :class:`BackendUnavailable`. ``last_kernel_ms`` holds the device
    time of the most recent integration kernel.
    """

    name: Final[str] = "hip"

    def __init__(self, device: int = 0) -> None:
        try:
            L = _lib.lib()
        except _lib.LibraryMissing as exc:
            raise BackendUnavailable(f"backend 'hip' is unavailable: {exc}") from exc
        count = L.kr_device_count()
        if count <= 0:
            raise BackendUnavailable("backend 'hip' is unavailable: no HIP device is visible")
        try:
            self.handle = _lib.Handle(device)
        except (_lib.KrError, ValueError) as exc:
            raise BackendUnavailable(f"backend 'hip' is unavailable: {exc}") from exc
        self.device = device
        self.last_kernel_ms = float("nan")

    def integrate_batch(
        self,
        st: Spacetime,
        Y0: ArrayLike,
        integ: IntegratorOptions,
        term: TerminationOptions,
        events: EventOptions | None = None,
    ) -> BatchResult:
        """Integrate the states ``Y0`` (N, 8); the fields of numba_backend.integrate_batch_numba."""
        if integ.method == "dop853":
            raise ValueError("method 'dop853' is available in integrate() only; the hip backend supports 'rk4' and 'rk45'")
        t0 = time.perf_counter()
        y0 = _validate_batch(Y0)
        rows = ResultRows(y0.shape[0], integ.np_dtype, events is not None and events.disk_plane)
        self.integrate_into(st, y0, integ, term, events, rows)
        return rows.result(time.perf_counter() - t0)

    def integrate_into(self, st: Spacetime, Y0: ArrayLike, integ: IntegratorOptions, term: TerminationOptions,
                       events: EventOptions | None, rows: ResultRows) -> None:
        """:meth:`integrate_batch` writing into the given row store (as many rows as ``Y0``)."""
        if integ.method == "dop853":
            raise ValueError("method 'dop853' is available in integrate() only; the hip backend supports 'rk4' and 'rk45'")
        y0 = _validate_batch(Y0)
        n = y0.shape[0]
        disk = events is not None and events.disk_plane
        if rows.n != n or rows.Y.dtype != integ.np_dtype or (disk and rows.event_Y is None):
            raise ValueError("the row store does not match the batch (rows, working dtype or event arrays)")
        params = kernel_params(st, integ, term, events)
        outputs = output_struct(rows)
        ms = ctypes.c_double(0.0)
        with self.handle._lock:
            _lib.check(
                _lib.lib().kr_integrate_batch(
                    self.handle.ptr, ctypes.byref(params), y0.ctypes.data, n, ctypes.byref(outputs),
                    ctypes.byref(ms),
                ),
                "kr_integrate_batch",
            )
        self.last_kernel_ms = ms.value

    def warm_up(self, st: Spacetime, integ: IntegratorOptions, term: TerminationOptions) -> float:
        """One short ray (``lambda_max = step_size``); returns the seconds it took (numba_backend.py:239-253)."""
        y0 = np.array([[0.0, 50.0, 0.5 * np.pi, 0.0, -1.0, -1.0, 0.0, 0.0]])
        short = IntegratorOptions(
            method=integ.method, rtol=integ.rtol, atol=integ.atol, step_size=integ.step_size,
            max_steps=integ.max_steps, lambda_max=integ.step_size, h_min=integ.h_min, dtype=integ.dtype,
        )
        t0 = time.perf_counter()
        self.integrate_batch(st, y0, short, term)
        return time.perf_counter() - t0

    def __repr__(self) -> str:
        return f"HipBackend(device={self.device})"


_HIP_CACHE: dict[int, HipBackend] = {}


def _hip_backend(device: int = 0) -> HipBackend:
    if device not in _HIP_CACHE:
        _HIP_CACHE[device] = HipBackend(device)
    return _HIP_CACHE[device]


def get_backend(name: str) -> Backend:
    """Return the backend called ``name`` (``"hip"`` here); see the module docstring."""
    if name == "hip":
        return _hip_backend()
    if name in ("numpy", "numba"):
        raise BackendUnavailable(REFERENCE_ONLY_MESSAGE.format(name=name))
    if name == "cuda":
        raise BackendUnavailable(CUDA_UNAVAILABLE_MESSAGE)
    raise ValueError(f"unknown backend {name!r}; expected one of {sorted(SUPPORTED_BACKENDS)}")


def available_backends() -> dict[str, bool]:
    """``{name: usable here}`` in :data:`BACKEND_NAMES` order."""
    try:
        _hip_backend()
        hip_ok = True
    except BackendUnavailable:
        hip_ok = False
    return {"numpy": False, "numba": False, "cuda": False, "hip": hip_ok}
