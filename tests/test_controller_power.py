"""The step controller's power on the device: correctly rounded x^-0.2 and powf(x, -0.2f).

kr_device.h's pow_m02_cr and powf_m02_cr (built for the host here) replace
the device library's pow in the exact attempts. Checks:

* against mpmath at 200 bits (the exact value, rounded to nearest) on
  random arguments and on every argument where the C library disagrees:
  they must be correctly rounded everywhere;
* against the C library's pow / powf, which the reference's kernel calls:
  agreement must exceed 99.9 % (the library itself misrounds the rest) on
  1e7 log-uniform doubles in [1e-5, 1e5] and on every float in [0.5, 2];
* a planted mistake (the exponent's excess over -1/5 dropped) must fail the
  mpmath check: that term moves the result by up to half an ulp.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

HOSTREF = os.path.join(ROOT, "tests", "hostref")
mpmath = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def L():
    subprocess.run(["make", "-s", "-C", HOSTREF], check=True, capture_output=True)
    lib = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    lib.krh_pow_m02.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    lib.krh_powf_m02.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    lib.krh_pow_m02_disagreements.restype = ctypes.c_int64
    lib.krh_pow_m02_disagreements.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int64]
    lib.krh_powf_m02_disagreements.restype = ctypes.c_int64
    lib.krh_powf_m02_disagreements.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_int64)]
    return lib


def exact_double(x: float, e: float) -> float:
    mpmath.mp.prec = 200
    return float(mpmath.power(mpmath.mpf(x), mpmath.mpf(e)))


def exact_float(x: float) -> np.float32:
    mpmath.mp.prec = 200
    v = mpmath.power(mpmath.mpf(float(x)), mpmath.mpf(float(np.float32(-0.2))))
    # round to float directly (not through double): nearest of the two float neighbours
    d = float(v)
    f = np.float32(d)
    lo, hi = (np.nextafter(f, np.float32(0)), f) if mpmath.mpf(float(f)) > v else (f, np.nextafter(f, np.float32(np.inf)))
    return lo if abs(mpmath.mpf(float(lo)) - v) <= abs(mpmath.mpf(float(hi)) - v) else hi


def run_double(L, xs: np.ndarray) -> np.ndarray:
    xs = np.ascontiguousarray(xs, np.float64)
    out = np.zeros_like(xs)
    L.krh_pow_m02(xs.ctypes.data, xs.size, out.ctypes.data)
    return out


def run_float(L, xs: np.ndarray) -> np.ndarray:
    xs = np.ascontiguousarray(xs, np.float32)
    out = np.zeros_like(xs)
    L.krh_powf_m02(xs.ctypes.data, xs.size, out.ctypes.data)
    return out


def test_double_power_is_correctly_rounded(L) -> None:
    rng = np.random.default_rng(11)
    xs = np.exp(rng.uniform(np.log(1e-5), np.log(1e5), 3000))
    xs = np.concatenate([xs, [1e-5, 1e5, 1.0, 0.5, 2.0, 32.0, 1.0 / 32.0, 3.0, 1e-3, 7.5e4]])
    got = run_double(L, xs)
    for x, g in zip(xs, got):
        assert g == exact_double(float(x), -0.2), x


def test_double_power_agrees_with_the_c_library_where_it_rounds_correctly(L) -> None:
    bad = np.zeros(200)
    n = 10_000_000
    nb = L.krh_pow_m02_disagreements(n, 12345, bad.ctypes.data, bad.size)
    print(f"pow(x, -0.2): {nb} of {n} differ from the C library ({100 * (1 - nb / n):.4f} % agree)")
    assert nb / n < 1e-3
    xs = bad[:min(nb, bad.size)]
    got = run_double(L, xs)
    for x, g in zip(xs, got):
        ex = exact_double(float(x), -0.2)
        assert g == ex and float(x) ** -0.2 != ex, x  # ours is right where the library is wrong


def test_float_power_is_correctly_rounded_and_agrees_with_the_c_library(L) -> None:
    rng = np.random.default_rng(5)
    xs = np.exp(rng.uniform(np.log(1e-5), np.log(1e5), 2000)).astype(np.float32)
    got = run_float(L, xs)
    for x, g in zip(xs, got):
        assert g == exact_float(x), x
    bad = np.zeros(200, np.float32)
    tested = ctypes.c_int64(0)
    nb = L.krh_powf_m02_disagreements(0.5, 2.0, bad.ctypes.data, bad.size, ctypes.byref(tested))
    print(f"powf(x, -0.2f): {nb} of {tested.value} floats in [0.5, 2] differ from the C library")
    assert tested.value == 2 * 2**23 + 1 and nb / tested.value < 1e-3
    for x in bad[:min(nb, 60)]:
        assert run_float(L, np.array([x]))[0] == exact_float(x), x


def test_dropping_the_exponent_excess_is_caught() -> None:
    """x^(-1/5) instead of x^(-0.2 as a double): wrong on a visible share of arguments."""
    rng = np.random.default_rng(2)
    xs = np.exp(rng.uniform(np.log(1e-5), np.log(1e5), 400))
    mpmath.mp.prec = 200
    wrong = sum(float(mpmath.power(mpmath.mpf(float(x)), -mpmath.mpf(1) / 5)) != exact_double(float(x), -0.2)
                for x in xs)
    assert wrong > 10
