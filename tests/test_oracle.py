"""The CPU oracle is pinned before it is trusted.

Pins, each chosen so that a plausible slip in the restatement fails one:

* golden vectors written by the reference itself (tests/golden/make_golden.py):
  the Numba kernel source run interpreted must agree bit for bit on every
  BatchResult field, the NumPy reference on the mask and step counts;
* closed forms: the Schwarzschild mask with Bardeen initial conditions is
  exactly b^2 = xi^2 + eta < 27 M^2, the DP5 tableau equals SciPy's RK45,
  the error norm and step factor reproduce the reference's known answers,
  the inverse metric inverts the covariant Kerr metric and its derivatives
  are the derivatives (finite differences, SymPy gradient of H);
* invariants: beta -> -beta mirror symmetry, E and L_z exactly conserved,
  the initial states are null;
* a planted mistake (one tableau coefficient moved by one ulp, inside the
  oracle) is caught by the golden comparison.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from conftest import ns

INTEG = dict(method="rk45", step_size=0.01, max_steps=100_000, lambda_max=1e4, h_min=1e-12, dtype="float64", h_max=None)
TERM = ns(horizon_epsilon=1e-6, escape_radius=1000.0)
FIELDS = ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift")


def integ(**kw):
    return ns(**{**INTEG, **kw})


def test_schwarzschild_64_oracle_is_bitwise_the_reference_kernel(oracle, golden1) -> None:
    Y0 = oracle.bardeen_image(1.0, 0.0, 1000.0, 60.0, 10.0, 64)
    assert np.array_equal(Y0, golden1["Y0"]), "Bardeen initial states differ from rayset._off_equatorial_state"
    r = oracle.integrate_batch(ns(mass=1.0, spin=0.0), Y0, integ(rtol=1e-8, atol=1e-10), TERM)
    for name in FIELDS:
        assert np.array_equal(getattr(r, name), golden1[f"numba_{name}"]), name
    # The reference's own two backends differ here: NumPy's vectorised power (SVML on AVX-512)
    # is not glibc's pow, so 2 of 4096 rays take one more accepted step on the NumPy path.
    assert np.array_equal(r.state, golden1["numpy_state"])
    differ = r.n_steps != golden1["numpy_n_steps"]
    assert np.count_nonzero(differ) <= 4 and np.all(np.abs(r.n_steps - golden1["numpy_n_steps"]) <= 1)
    assert np.all(r.max_energy_drift == 0.0) and np.all(r.max_lz_drift == 0.0)


def test_schwarzschild_64_mask_is_the_closed_form(oracle) -> None:
    """a = 0: captured <=> xi^2 + eta = alpha^2 + beta^2 < 27 M^2; nearest centres are 0.019 M away."""
    Y0 = oracle.bardeen_image(1.0, 0.0, 1000.0, 60.0, 10.0, 64)
    r = oracle.integrate_batch(ns(mass=1.0, spin=0.0), Y0, integ(rtol=1e-8, atol=1e-10), TERM)
    c = oracle.pixel_centres(10.0, 64)
    alpha, beta = np.meshgrid(c, -c)
    b2 = (alpha * alpha + beta * beta).ravel()
    assert np.min(np.abs(np.sqrt(b2) - math.sqrt(27.0))) > 0.018
    assert set(np.unique(r.state)) == {1, 2}
    assert np.array_equal(r.state == 2, b2 < 27.0)


def test_kerr_512_oracle_matches_reference_golden(oracle, golden2) -> None:
    """a = 0.9, i = 60 deg, 512^2, rtol 1e-6: bitwise on the edge band + sample, mask on every pixel."""
    import hashlib

    st = ns(mass=1.0, spin=0.9)
    Y0 = oracle.bardeen_image(1.0, 0.9, 1000.0, 60.0, 10.0, 512)
    assert hashlib.sha256(Y0.tobytes()).hexdigest() == str(golden2["Y0_sha256"])
    pick = golden2["pick"]
    assert np.array_equal(Y0[pick], golden2["Y0_pick"])
    full = oracle.integrate_batch(st, Y0, integ(rtol=1e-6, atol=1e-8), TERM, dynamic=True)
    assert np.array_equal(full.state, golden2["numpy_state"].astype(np.int64))
    assert np.array_equal(full.n_steps, golden2["numpy_n_steps"].astype(np.int64))
    assert np.array_equal(full.n_rejected, golden2["numpy_n_rejected"].astype(np.int64))
    for name in FIELDS:
        assert np.array_equal(getattr(full, name)[pick], golden2[f"numba_{name}"]), name
    mask = (full.state == 2).reshape(512, 512)
    assert np.array_equal(mask[::-1, :], mask), "beta -> -beta mirror symmetry"


def test_planted_tableau_error_is_caught(oracle, golden1) -> None:
    """A one-ulp change of one DP5 coefficient inside the oracle breaks the bitwise golden match."""
    Y0 = golden1["Y0"][::37]
    st, it = ns(mass=1.0, spin=0.0), integ(rtol=1e-8, atol=1e-10)
    good = oracle.integrate_batch(st, Y0, it, TERM)
    assert np.array_equal(good.Y, golden1["numba_Y"][::37])
    tab = oracle.tableaus
    try:
        def bad_tableaus():
            a, e, b, c = tab()
            a = a.copy()
            a[4, 2] = np.nextafter(a[4, 2], 10.0)
            return a, e, b, c

        oracle.tableaus = bad_tableaus
        bad = oracle.integrate_batch(st, Y0, it, TERM)
    finally:
        oracle.tableaus = tab
    assert not np.array_equal(bad.Y, golden1["numba_Y"][::37])


def test_tableau_matches_scipy_rk45(oracle) -> None:
    from scipy.integrate import RK45

    a, e, b, c = oracle.tableaus()
    assert np.allclose(a[:6, :5], RK45.A, rtol=0.0, atol=1e-15)
    assert np.allclose(e, RK45.E, rtol=0.0, atol=1e-15)
    assert np.allclose(a[6, :6], RK45.B, rtol=0.0, atol=1e-15)
    assert np.allclose(a.sum(axis=1)[1:], np.asarray(RK45.C[1:].tolist() + [1.0]), atol=1e-15)
    assert abs(b.sum() - 1.0) < 1e-15 and tuple(c) == (0.0, 0.5, 0.5, 1.0)


def test_error_norm_and_step_factor_known_answers(oracle) -> None:
    """The values of tests/test_integrators.py:64-76 of the reference."""
    import ctypes

    L = oracle.lib()
    _, consts, _, _ = oracle.kernel_constants(ns(mass=1.0, spin=0.0), integ(rtol=1e-6, atol=1e-8), TERM, None)
    y = np.zeros(8)
    yn = np.zeros(8)
    err = np.zeros(8)
    y[0], y[1] = 1.0, -2.0
    yn[0], yn[1] = 1.5, -1.0
    err[0], err[1] = 1e-6, 2e-6
    sc = 1e-8 + 1e-6 * np.array([1.5, 2.0])
    expected = math.sqrt(np.sum((err[:2] / sc) ** 2) / 8.0)
    got = L.kro_error_norm(consts.ctypes.data, err.ctypes.data, y.ctypes.data, yn.ctypes.data)
    assert abs(got - expected) < 1e-15 * expected
    f = lambda en, rb=0: L.kro_step_factor(consts.ctypes.data, ctypes.c_double(en), rb)  # noqa: E731
    assert f(1.0) == pytest.approx(0.9)
    assert f(0.0) == 5.0 and f(1e12) == 0.2 and f(1e-12) == 5.0
    assert f(1e-12, 1) == 1.0
    assert f(2.0**5) == pytest.approx(0.45)


def _covariant_metric(m: float, a: float, r: float, th: float) -> np.ndarray:
    """MTW (1973) eq. 33.2 in Boyer-Lindquist coordinates."""
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    sig = r * r + a * a * c2
    dlt = r * r - 2 * m * r + a * a
    g = np.zeros((4, 4))
    g[0, 0] = -(1 - 2 * m * r / sig)
    g[0, 3] = g[3, 0] = -2 * m * a * r * s2 / sig
    g[1, 1] = sig / dlt
    g[2, 2] = sig
    g[3, 3] = (r * r + a * a + 2 * m * a * a * r * s2 / sig) * s2
    return g


@pytest.mark.parametrize("spin", [0.0, 0.9, -0.7])
def test_inverse_metric_and_derivatives(oracle, spin: float) -> None:
    L = oracle.lib()
    geom, _, _, _ = oracle.kernel_constants(ns(mass=1.0, spin=spin), integ(rtol=1e-6, atol=1e-8), TERM, None)
    rng = np.random.default_rng(7)
    out = np.empty(15)

    def terms(r, th):
        L.kro_inverse_metric_terms(geom.ctypes.data, r, th, out.ctypes.data)
        return out.copy()

    for _ in range(50):
        r, th = rng.uniform(2.0, 40.0), rng.uniform(0.1, math.pi - 0.1)
        t = terms(r, th)
        ginv = np.zeros((4, 4))
        ginv[0, 0], ginv[0, 3], ginv[3, 0], ginv[1, 1], ginv[2, 2], ginv[3, 3] = t[0], t[1], t[1], t[2], t[3], t[4]
        assert np.allclose(ginv @ _covariant_metric(1.0, spin, r, th), np.eye(4), atol=1e-12)
        h = 1e-6
        d_r = (terms(r + h, th)[:5] - terms(r - h, th)[:5]) / (2 * h)
        d_th = (terms(r, th + h)[:5] - terms(r, th - h)[:5]) / (2 * h)
        assert np.allclose(t[5:10], d_r, rtol=1e-6, atol=1e-9)
        assert np.allclose(t[10:15], d_th, rtol=1e-6, atol=1e-9)


def test_rhs_is_hamiltons_equations_sympy(oracle) -> None:
    """f = (dH/dp, -dH/dx) for H = 1/2 g^{mu nu} p_mu p_nu (BPT 1972 eq. 2.1 definitions), via SymPy."""
    import sympy as sp

    r, th, pt, pr, pth, pph = sp.symbols("r th pt pr pth pph", real=True)
    m, a = 1, sp.Rational(9, 10)
    sig = r**2 + a**2 * sp.cos(th) ** 2
    dlt = r**2 - 2 * m * r + a**2
    big_a = (r**2 + a**2) ** 2 - a**2 * dlt * sp.sin(th) ** 2
    H = sp.Rational(1, 2) * (
        -big_a / (sig * dlt) * pt**2 + 2 * (-2 * m * a * r / (sig * dlt)) * pt * pph + dlt / sig * pr**2
        + pth**2 / sig + (dlt - a**2 * sp.sin(th) ** 2) / (sig * dlt * sp.sin(th) ** 2) * pph**2
    )
    exprs = [sp.diff(H, pt), sp.diff(H, pr), sp.diff(H, pth), sp.diff(H, pph), -sp.diff(H, r), -sp.diff(H, th)]
    fn = sp.lambdify((r, th, pt, pr, pth, pph), exprs, "math")
    L = oracle.lib()
    geom, consts, _, _ = oracle.kernel_constants(ns(mass=1.0, spin=0.9), integ(rtol=1e-6, atol=1e-8), TERM, None)
    rng = np.random.default_rng(3)
    f = np.empty(8)
    for _ in range(40):
        y = np.array([0.0, rng.uniform(2.0, 30.0), rng.uniform(0.2, 2.9), 0.0, -1.0,
                      rng.uniform(-2, 2), rng.uniform(-3, 3), rng.uniform(-4, 4)])
        L.kro_rhs(geom.ctypes.data, consts.ctypes.data, y.ctypes.data, f.ctypes.data)
        ref = fn(y[1], y[2], y[4], y[5], y[6], y[7])
        got = [f[0], f[1], f[2], f[3], f[5], f[6]]
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-12)
        assert f[4] == 0.0 and f[7] == 0.0


def test_initial_states_are_null_and_carry_bardeen_constants(oracle) -> None:
    L = oracle.lib()
    geom, consts, _, _ = oracle.kernel_constants(ns(mass=1.0, spin=0.9), integ(rtol=1e-6, atol=1e-8), TERM, None)
    c = oracle.pixel_centres(10.0, 16)
    alpha, beta = np.meshgrid(c, -c)
    Y0 = oracle.bardeen_image(1.0, 0.9, 1000.0, 60.0, 10.0, 16)
    th = math.radians(60.0)
    assert np.allclose(Y0[:, 7], -alpha.ravel() * math.sin(th), rtol=0, atol=0)
    # Theta(theta_o) = eta - cos^2 (xi^2/sin^2 - a^2) = beta^2 (Bardeen's beta = p_theta / E at the observer).
    assert np.allclose(Y0[:, 6] ** 2, beta.ravel() ** 2, rtol=1e-12, atol=1e-12)
    assert np.all(np.sign(Y0[:, 6]) == -np.sign(beta.ravel()))
    f = np.empty(8)
    for y in Y0:
        L.kro_rhs(geom.ctypes.data, consts.ctypes.data, y.ctypes.data, f.ctypes.data)
        h = 0.5 * (y[4] * f[0] + y[5] * f[1] + y[6] * f[2] + y[7] * f[3])
        assert abs(h) < 1e-9


def test_classification_priority(oracle) -> None:
    """Rule order of classification.py:1-29 (tests/test_classification.py:109-118 of the reference)."""
    L = oracle.lib()
    _, consts, params, ipar = oracle.kernel_constants(ns(mass=1.0, spin=0.9), integ(rtol=1e-6, atol=1e-8), TERM, None)
    r_cap = consts[12]

    def cls(y, f, lam=0.0, n=0, h=0.01, check=1):
        y, f = np.asarray(y, float), np.asarray(f, float)
        return L.kro_classify(consts.ctypes.data, params.ctypes.data, ipar.ctypes.data, y.ctypes.data,
                              f.ctypes.data, lam, n, h, check)

    base = [0, 10.0, 1.0, 0, -1, 0.1, 0.1, 2.0]
    f_out = [1, 1.0, 0, 0, 0, 0, 0, 0]
    assert cls(base, f_out) == 0
    assert cls([0, np.nan] + base[2:], f_out) == 4
    assert cls(base, f_out, h=1e-13) == 4 and cls(base, f_out, h=1e-13, check=0) == 0
    assert cls([0, -1.0] + base[2:], f_out) == 5
    assert cls([0, 10.0, -2e-9] + base[3:], f_out) == 5
    assert cls([0, 10.0, 1e-13] + base[3:], f_out) == 5
    assert cls([0, 10.0, 1e-13, 0, -1, 0.1, 0.1, 0.0], f_out) == 0  # L_z = 0 may touch the axis
    assert cls([0, r_cap] + base[2:], f_out) == 2
    assert cls([0, 1000.0] + base[2:], f_out) == 1
    assert cls([0, 1000.0] + base[2:], [1, -1.0, 0, 0, 0, 0, 0, 0]) == 0  # inward at r_esc: not escaped
    assert cls(base, f_out, lam=1e4) == 3 and cls(base, f_out, n=100_000) == 3
    assert cls([0, r_cap, -2e-9] + base[3:], f_out, h=1e-13) == 4  # failure outranks everything
    assert cls([0, -1.0] + base[2:], f_out, lam=1e4) == 5  # domain outranks capture/budget
