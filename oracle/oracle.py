"""ctypes front end of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module; the product path never does. It packs the
kernel constants exactly as the reference's Numba host wrapper does
(``src/kerrray/raytracing/backends/numba_backend.py:87-128,167-207``) and
runs the C++ restatement in ``kr_oracle.cpp``.

Inputs are duck-typed: any object with the attribute names of the
reference's ``Spacetime`` (``mass``, ``spin``), ``IntegratorOptions``
(``method``, ``rtol``, ``atol``, ``step_size``, ``max_steps``, ``lambda_max``,
``h_min``, ``dtype``, ``h_max``), ``TerminationOptions``
(``horizon_epsilon``, ``escape_radius``) and ``EventOptions``
(``disk_plane``, ``r_in``, ``r_out``) works.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
import time
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkr_oracle.so")

# Dormand-Prince 5(4) and RK4 tableaus (tableaus.py:59-91; Dormand & Prince 1980;
# Hairer, Norsett & Wanner 1993 Table 5.2), evaluated in float64 as Python does.
RK4_C = (0.0, 0.5, 0.5, 1.0)
RK4_B = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)
DP_A = (
    (),
    (1.0 / 5.0,),
    (3.0 / 40.0, 9.0 / 40.0),
    (44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0),
    (19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0),
    (9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0),
    (35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0),
)
DP_B = (35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0)
DP_B_HAT = (5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0)
DP_E = tuple(bh - b for bh, b in zip(DP_B_HAT, DP_B))
SAFETY, FACTOR_MIN, FACTOR_MAX, ERROR_EXPONENT = 0.9, 0.2, 5.0, -1.0 / 5.0
THETA_DOMAIN_TOLERANCE = 1e-9
AXIS_SIN_TOLERANCE = 1e-12

_lib = None


def build() -> str:
    """Compile ``libkr_oracle.so`` with ``make`` (idempotent) and return its path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        vp = ctypes.c_void_p
        L.kro_integrate_rays.restype = ctypes.c_int
        L.kro_integrate_rays.argtypes = [ctypes.c_int] + [vp] * 27 + [ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.kro_bardeen_image.restype = ctypes.c_int
        L.kro_bardeen_image.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int64, vp]
        L.kro_bardeen_state.restype = ctypes.c_int
        L.kro_bardeen_state.argtypes = [ctypes.c_double] * 6 + [ctypes.c_int, vp]
        L.kro_pixel_centres.restype = None
        L.kro_pixel_centres.argtypes = [ctypes.c_double, ctypes.c_int64, vp]
        L.kro_inverse_metric_terms.restype = None
        L.kro_inverse_metric_terms.argtypes = [vp, ctypes.c_double, ctypes.c_double, vp]
        L.kro_rhs.restype = None
        L.kro_rhs.argtypes = [vp, vp, vp, vp]
        L.kro_error_norm.restype = ctypes.c_double
        L.kro_error_norm.argtypes = [vp, vp, vp, vp]
        L.kro_step_factor.restype = ctypes.c_double
        L.kro_step_factor.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.kro_dp5_step.restype = None
        L.kro_dp5_step.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, vp, vp, vp, vp]
        L.kro_classify.restype = ctypes.c_int
        L.kro_classify.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_int64, ctypes.c_double, ctypes.c_int]
        L.kro_disk_shade.restype = ctypes.c_int
        L.kro_disk_shade.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int64, vp, vp]
        L.kro_zamo_frame.restype = ctypes.c_int
        L.kro_zamo_frame.argtypes = [ctypes.c_double] * 4 + [vp] * 3
        L.kro_zamo_states.restype = ctypes.c_int
        L.kro_zamo_states.argtypes = [ctypes.c_double] * 5 + [vp, vp, ctypes.c_int64, ctypes.c_int, vp]
        L.kro_escape_variational.restype = ctypes.c_int
        L.kro_escape_variational.argtypes = [ctypes.c_double] * 4 + [vp, vp, ctypes.c_int64] + [vp] * 13 + [ctypes.c_int]
        del dp, ip
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if a is None else a.ctypes.data)


def outer_horizon(mass: float, spin: float) -> float:
    """``r_+ = M + sqrt(M^2 - a^2)`` (geometry/horizons.py:33-42)."""
    a = spin * mass
    return mass + math.sqrt(mass * mass - a * a)


def max_attempts(integ) -> int:
    """Defensive loop guard (geodesics/integrators.py:151-162)."""
    h_max = max(integ.lambda_max, integ.step_size)
    rej = (math.log(h_max / integ.h_min) + integ.max_steps * math.log(FACTOR_MAX)) / math.log(1.0 / SAFETY)
    return integ.max_steps + int(math.ceil(rej)) + 1


def kernel_constants(st, integ, term, events):
    """``(geom, consts, params, ipar)`` as numba_backend.py:87-115 builds them (consts kept in float64)."""
    m = float(st.mass)
    a = float(st.spin) * m
    geom = np.array([m, a, a**2, -2.0 * m * a, 2.0 * m * a, 2.0 * m, 2.0 * a**2], dtype=np.float64)
    disk = events is not None and bool(events.disk_plane)
    r_in = float(events.r_in) if events is not None else 0.0
    r_out = float(events.r_out) if events is not None else 0.0
    consts = np.array(
        [
            0.0, 1.0, 2.0, -0.5, 8.0, SAFETY, FACTOR_MIN, FACTOR_MAX, ERROR_EXPONENT,
            integ.rtol, integ.atol, integ.h_min, outer_horizon(m, float(st.spin)) + term.horizon_epsilon,
            term.escape_radius, -THETA_DOMAIN_TOLERANCE, np.pi + THETA_DOMAIN_TOLERANCE,
            AXIS_SIN_TOLERANCE, integ.step_size, r_in, r_out, 0.5 * np.pi,
        ],
        dtype=np.float64,
    )
    h_max = np.inf if integ.h_max is None else integ.h_max
    params = np.array([integ.lambda_max, 0.5 * np.pi, h_max], dtype=np.float64)
    ipar = np.array([integ.max_steps, max_attempts(integ), int(integ.method == "rk45"), int(disk)], dtype=np.int64)
    return geom, consts, params, ipar


def tableaus():
    """Dense DP ``A`` (7x7), ``E`` (7,), RK4 ``B``, ``C`` in float64 (numba_backend.py:118-128)."""
    tab_a = np.zeros((7, 7), dtype=np.float64)
    for i, row in enumerate(DP_A):
        tab_a[i, : len(row)] = row
    return tab_a, np.asarray(DP_E, np.float64), np.asarray(RK4_B, np.float64), np.asarray(RK4_C, np.float64)


@dataclass
class OracleResult:
    """The reference ``BatchResult`` fields (photons/trajectories.py:100-128)."""

    Y: np.ndarray
    state: np.ndarray
    n_steps: np.ndarray
    lam: np.ndarray
    max_null_error: np.ndarray
    max_energy_drift: np.ndarray
    max_lz_drift: np.ndarray
    max_carter_drift: np.ndarray
    runtime_s: float
    event_Y: np.ndarray | None
    event_hit: np.ndarray | None
    n_rejected: np.ndarray


def integrate_batch(st, Y0, integ, term, events=None, *, n_threads: int = 0, dynamic: bool = False) -> OracleResult:
    """Run the restated kernel over the rows of ``Y0`` (N, 8)."""
    if integ.method == "dop853":
        raise ValueError("method 'dop853' is not supported by the batched kernel")
    y0 = np.array(Y0, dtype=np.float64)
    if y0.ndim != 2 or y0.shape[1] != 8:
        raise ValueError(f"Y0 must have shape (N, 8), got {y0.shape}")
    if not np.all(np.isfinite(y0)):
        raise ValueError("Y0 must be finite")
    n = y0.shape[0]
    dtype = np.dtype(integ.dtype)
    geom, consts, params, ipar = kernel_constants(st, integ, term, events)
    tab_a, tab_e, rk4_b, rk4_c = tableaus()
    Y = np.ascontiguousarray(y0.astype(dtype))
    F = np.empty_like(Y)
    lam = np.zeros(n)
    n_steps = np.zeros(n, np.int64)
    n_rej = np.zeros(n, np.int64)
    state = np.zeros(n, np.int64)
    maxima = [np.zeros(n) for _ in range(4)]
    # numba_backend.py:174-178 (photons/constants.py:65-90,163-171)
    a = float(st.spin) * float(st.mass)
    e0 = np.ascontiguousarray(-y0[:, 4])
    lz0 = np.ascontiguousarray(y0[:, 7])
    th, p_th = y0[:, 2], y0[:, 6]
    s2 = np.sin(th) ** 2
    lz2_over_s2 = np.divide(lz0 * lz0, s2, out=np.zeros(n), where=lz0 != 0)
    q0 = np.ascontiguousarray(p_th * p_th + np.cos(th) ** 2 * (lz2_over_s2 - a**2 * e0 * e0))
    fe = np.abs(e0)
    fe = np.where(fe > 0.0, fe, 1.0)
    fl = fe * float(st.mass)
    fq = fe * fe * float(st.mass) * float(st.mass)
    fe2 = fe**2
    disk = bool(ipar[3])
    event_Y = np.zeros((n if disk else 0, 8), dtype=dtype)
    event_hit = np.zeros(n if disk else 0, dtype=np.uint8)
    t0 = time.perf_counter()
    rc = lib().kro_integrate_rays(
        0 if dtype == np.float64 else 1,
        *(_ptr(x) for x in (geom, consts, params, ipar, tab_a, tab_e, rk4_b, rk4_c, Y, F, lam, n_steps, n_rej, state,
                            *maxima, e0, lz0, q0, fe, fl, fq, fe2, event_Y, event_hit)),
        n, n_threads, 1 if dynamic else 0,
    )
    if rc != 0:
        raise ValueError("oracle rejected its inputs")
    return OracleResult(
        Y=Y, state=state, n_steps=n_steps, lam=lam, max_null_error=maxima[0], max_energy_drift=maxima[1],
        max_lz_drift=maxima[2], max_carter_drift=maxima[3], runtime_s=time.perf_counter() - t0,
        event_Y=event_Y if disk else None, event_hit=event_hit.astype(bool) if disk else None, n_rejected=n_rej,
    )


def pixel_centres(fov: float, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().kro_pixel_centres(fov, n, _ptr(out))
    return out


def bardeen_image(mass: float, spin: float, r_o: float, inclination_deg: float, fov: float, n: int) -> np.ndarray:
    """Row-major (n*n, 8) Bardeen initial states of an n x n image (restated in C)."""
    Y = np.empty((n * n, 8))
    rc = lib().kro_bardeen_image(mass, spin * mass, r_o, math.radians(inclination_deg), fov, n, _ptr(Y))
    if rc != 0:
        raise ValueError("negative radicand in the Bardeen initial conditions")
    return Y


def bardeen_points(mass: float, spin: float, r_o: float, inclination_deg: float, alpha, beta) -> np.ndarray:
    """Bardeen initial states for arbitrary image points (sign_theta = -1 for beta >= 0)."""
    alpha = np.asarray(alpha, np.float64).ravel()
    beta = np.asarray(beta, np.float64).ravel()
    Y = np.empty((alpha.size, 8))
    th = math.radians(inclination_deg)
    L = lib()
    for k in range(alpha.size):
        yk = np.empty(8)
        if L.kro_bardeen_state(mass, spin * mass, r_o, th, float(alpha[k]), float(beta[k]),
                               -1 if beta[k] >= 0.0 else 1, _ptr(yk)):
            raise ValueError("negative radicand in the Bardeen initial conditions")
        Y[k] = yk
    return Y


@dataclass
class EscapeJacobian:
    state: np.ndarray  # (n,) termination codes of the exact integration
    theta: np.ndarray  # (n,) final theta
    phi: np.ndarray  # (n,) final phi
    jac: np.ndarray  # (n, 2, 4): d(theta, phi)/d(alpha, beta, spin, theta_o); NaN unless escaped
    replay_err: np.ndarray  # (n,) relative difference of the variational pass's values from the exact states


def escape_variational(st, theta_o: float, r_o: float, alpha, beta, integ, term, *, n_threads: int = 0) -> EscapeJacobian:
    """Forward-mode variational integration along the accepted steps (see kr_oracle.cpp).

    ``theta_o`` in radians; derivatives are with respect to
    ``(alpha, beta, spin, theta_o)``.
    """
    alpha = np.ascontiguousarray(alpha, np.float64).ravel()
    beta = np.ascontiguousarray(beta, np.float64).ravel()
    n = alpha.size
    geom, consts, params, ipar = kernel_constants(st, integ, term, None)
    tab_a, tab_e, rk4_b, rk4_c = tableaus()
    state = np.empty(n, np.int64)
    theta, phi, err = np.empty(n), np.empty(n), np.empty(n)
    jac = np.empty((n, 2, 4))
    rc = lib().kro_escape_variational(
        float(st.mass), float(st.spin), float(r_o), float(theta_o), _ptr(alpha), _ptr(beta), n, _ptr(geom),
        _ptr(consts), _ptr(params), _ptr(ipar), _ptr(tab_a), _ptr(tab_e), _ptr(rk4_b), _ptr(rk4_c), _ptr(state),
        _ptr(theta), _ptr(phi), _ptr(jac), _ptr(err), n_threads)
    if rc == 2:
        raise ValueError("the variational integration needs adaptive DP5 without events")
    if rc != 0:
        raise ValueError("negative radicand in the Bardeen initial conditions")
    return EscapeJacobian(state, theta, phi, jac, err)


@dataclass
class ZamoFrame:
    tetrad: np.ndarray
    metric: np.ndarray
    inverse_metric: np.ndarray


def zamo_frame(mass: float, spin: float, r: float, theta: float) -> ZamoFrame:
    """ZAMO tetrad rows, metric and inverse metric at (r, theta) (camera.py:225-262, metric.py:147-219)."""
    e, g, gi = np.empty((4, 4)), np.empty((4, 4)), np.empty((4, 4))
    if lib().kro_zamo_frame(mass, spin, r, theta, _ptr(e), _ptr(g), _ptr(gi)):
        raise ValueError("the ZAMO tetrad does not exist here")
    return ZamoFrame(e, g, gi)


def zamo_states(mass: float, spin: float, r_o: float, theta_o: float, phi_o: float, alpha, beta,
                reverse: bool = True) -> tuple[np.ndarray, int]:
    """``initial_states`` of a ZAMO camera at image points (camera.py:298-381): (N, 8) states and the flag bits
    (1 no frame, 2 not null, 4 outside the sky)."""
    a = np.ascontiguousarray(np.asarray(alpha, dtype=np.float64).ravel())
    b = np.ascontiguousarray(np.asarray(beta, dtype=np.float64).ravel())
    Y = np.empty((a.size, 8))
    bad = lib().kro_zamo_states(mass, spin, r_o, theta_o, phi_o, _ptr(a), _ptr(b), a.size, 1 if reverse else 0,
                                _ptr(Y))
    return Y, int(bad)


def disk_shade(mass: float, spin: float, r_in: float, r_out: float, emissivity_index: float, redshift_power: int,
               prograde: bool, y_hit) -> tuple[np.ndarray, np.ndarray]:
    """(g, I_obs) of disk crossings (redshift.py:63-151, accretion.py:180-196); ValueError where the reference raises."""
    y = np.ascontiguousarray(np.asarray(y_hit, dtype=np.float64).reshape(-1, 8))
    g, intensity = np.empty(len(y)), np.empty(len(y))
    if lib().kro_disk_shade(mass, spin, r_in, r_out, emissivity_index, int(redshift_power), 1 if prograde else 0,
                            _ptr(y), len(y), _ptr(g), _ptr(intensity)):
        raise ValueError("a disk crossing cannot be shaded")
    return g, intensity


def _bilinear(field: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Bilinear value of ``field`` at fractional (row v, column u), zero ring outside (boundary.py:95-112)."""
    n_rows, n_cols = field.shape
    padded = np.pad(field, 1)
    up, vp = u + 1.0, v + 1.0
    inside = (up >= 0.0) & (up <= n_cols) & (vp >= 0.0) & (vp <= n_rows)
    j0 = np.clip(np.floor(up).astype(np.intp), 0, n_cols)
    i0 = np.clip(np.floor(vp).astype(np.intp), 0, n_rows)
    fu = np.clip(up - j0, 0.0, 1.0)
    fv = np.clip(vp - i0, 0.0, 1.0)
    j1 = np.minimum(j0 + 1, n_cols + 1)
    i1 = np.minimum(i0 + 1, n_rows + 1)
    value = (padded[i0, j0] * (1.0 - fu) * (1.0 - fv) + padded[i0, j1] * fu * (1.0 - fv)
             + padded[i1, j0] * (1.0 - fu) * fv + padded[i1, j1] * fu * fv)
    return np.where(inside, value, 0.0)


def mask_edge(alpha, beta, captured, n_angles: int = 360, centre=None) -> tuple[np.ndarray, np.ndarray]:
    """``extract_boundary_from_mask`` (boundary.py:74-170): (angles, radii) of the 0.5 level set of the
    bilinearly interpolated mask, marched from ``centre`` (the captured centroid by default) in steps of
    1/8 pixel."""
    a = np.asarray(alpha, dtype=np.float64)
    b = np.asarray(beta, dtype=np.float64)
    mask = np.asarray(captured, dtype=bool)
    d = float(a[0, 1] - a[0, 0])
    a0, b0 = float(a[0, 0]), float(b[0, 0])
    if centre is None:
        ca, cb = float(a[mask].mean()), float(b[mask].mean())
    else:
        ca, cb = float(centre[0]), float(centre[1])
    field = mask.astype(np.float64)
    angles = 2.0 * np.pi * np.arange(n_angles, dtype=np.float64) / n_angles
    ds = 0.125 * d
    reach = math.hypot(a[0, -1] - a[0, 0], b[0, 0] - b[-1, 0]) + 2.0 * d
    s = ds * np.arange(int(math.ceil(reach / ds)) + 1, dtype=np.float64)
    x = ca + np.cos(angles)[:, None] * s[None, :]
    y = cb + np.sin(angles)[:, None] * s[None, :]
    f = _bilinear(field, (x - a0) / d, (b0 - y) / d)
    if np.any(f[:, 0] < 0.5):
        raise ValueError("the centre lies outside the captured region")
    k = np.argmax(f < 0.5, axis=1)
    rows = np.arange(n_angles)
    f_in, f_out = f[rows, k - 1], f[rows, k]
    return angles, s[k - 1] + ds * (f_in - 0.5) / (f_in - f_out)
