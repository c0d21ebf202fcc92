"""The device kernel source, compiled for the CPU, is bit-identical to the oracle.

tests/hostref builds kerrray_amd/csrc/kr_device.h with g++ against glibc: the
exact per-photon code the GPU runs (six-component state, reused sin/cos, the
same operation order), with the host's libm in place of the device's. Every
BatchResult field must then equal the oracle's (and the reference's golden
vectors) bit for bit, for every option the kernel supports. On the GPU the
only remaining difference is the device math library (sin, cos, pow round
differently from glibc in a few percent of calls), which the GPU tests bound.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT, options

from kerrray_amd.backends import HostBatch, kernel_params
from kerrray_amd.camera import BardeenCamera, bardeen_initial_states, bardeen_points
from kerrray_amd.geodesics import EventOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.shadow import kr_image

FIELDS = ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_energy_drift", "max_lz_drift",
          "max_carter_drift")
HOSTREF = os.path.join(ROOT, "tests", "hostref")


@pytest.fixture(scope="module")
def hostref():
    subprocess.run(["make", "-s", "-C", HOSTREF], check=True)
    L = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    L.krh_integrate.restype = ctypes.c_int
    L.krh_integrate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]
    L.krh_trace_image.restype = ctypes.c_int
    L.krh_trace_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return L


def run_y0(L, st, Y0, integ, term, events=None):
    Y0 = np.ascontiguousarray(Y0, dtype=np.float64)
    p = kernel_params(st, integ, term, events)
    out = HostBatch(len(Y0), integ.np_dtype, events is not None and events.disk_plane)
    assert L.krh_integrate(ctypes.addressof(p), Y0.ctypes.data, len(Y0), ctypes.addressof(out.outputs), 0) == 0
    return out.result(0.0)


def run_image(L, st, cam, integ, term):
    p = kernel_params(st, integ, term)
    img = kr_image(st, cam.radius, cam.inclination_deg, cam.fov, cam.resolution)
    out = HostBatch(cam.n_rays, integ.np_dtype, False)
    assert L.krh_trace_image(ctypes.addressof(p), ctypes.addressof(img), ctypes.addressof(out.outputs), 0) == 0
    return out.result(0.0)


def same(a, b, fields=FIELDS) -> None:
    for name in fields:
        x, y = getattr(a, name), getattr(b, name)
        if x is None and y is None:
            continue
        assert np.array_equal(np.asarray(x), np.asarray(y)), name


def test_schwarzschild_64_image_is_bitwise_the_reference(hostref, golden1) -> None:
    integ, term = options(rtol=1e-8, atol=1e-10)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=64)
    r = run_image(hostref, Spacetime(1.0, 0.0), cam, integ, term)
    for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift"):
        assert np.array_equal(getattr(r, name), golden1[f"numba_{name}"]), name


def test_kerr_512_image_is_bitwise_the_reference(hostref, golden2) -> None:
    integ, term = options(rtol=1e-6, atol=1e-8)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
    r = run_image(hostref, Spacetime(1.0, 0.9), cam, integ, term)
    assert np.array_equal(r.state, golden2["numpy_state"].astype(np.int64))
    pick = golden2["pick"]
    for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift"):
        assert np.array_equal(getattr(r, name)[pick], golden2[f"numba_{name}"]), name


@pytest.mark.parametrize("kw", [
    dict(rtol=1e-8, atol=1e-10),
    dict(rtol=1e-6, atol=1e-8, dtype="float32", horizon_epsilon=1e-3),
    dict(method="rk4", step_size=0.05, lambda_max=20.0, escape_radius=1e6),
    dict(rtol=1e-8, atol=1e-10, max_steps=30),
    dict(rtol=1e-8, atol=1e-10, lambda_max=500.0),
    dict(rtol=1e-8, atol=1e-10, h_max=0.5),
    dict(method="rk4", step_size=0.05, lambda_max=3000.0, dtype="float32", horizon_epsilon=1e-3),
])
def test_options_are_bitwise_the_oracle(hostref, oracle, kw) -> None:
    st = Spacetime(1.0, 0.9)
    integ, term = options(**kw)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=16), st)
    same(run_y0(hostref, st, Y0, integ, term), oracle.integrate_batch(st, Y0, integ, term))


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_disk_event_is_bitwise_the_oracle(hostref, oracle, dtype) -> None:
    st = Spacetime(1.0, 0.9)
    th = 1.2
    lz = np.array([-6.0, 3.0, 6.0, 9.0])
    Y0 = bardeen_points(st, 200.0, math.degrees(th), -lz / math.sin(th),
                        -np.sqrt(15.0 + 0.81 * math.cos(th) ** 2 - lz**2 / math.tan(th) ** 2))
    events = EventOptions(disk_plane=True, r_in=6.0, r_out=30.0)
    integ, far = options(rtol=1e-7, atol=1e-9, lambda_max=400.0, dtype=dtype, escape_radius=1e6)
    a = run_y0(hostref, st, Y0, integ, far, events)
    b = oracle.integrate_batch(st, Y0, integ, far, events=events)
    same(a, b, FIELDS + ("event_Y", "event_hit"))
    assert a.event_hit.all()


def test_schwarzschild_and_negative_spin_are_bitwise_the_oracle(hostref, oracle) -> None:
    for spin, incl in ((0.0, 89.0), (-0.7, 30.0), (0.99, 85.0)):
        st = Spacetime(1.0, spin)
        integ, term = options(rtol=1e-9, atol=1e-11)
        Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=incl, fov=8.0, resolution=8), st)
        same(run_y0(hostref, st, Y0, integ, term), oracle.integrate_batch(st, Y0, integ, term))


def test_slopes_held_outside_registers_give_the_same_bits(hostref, golden2) -> None:
    """dp5_step_stored (k1..k4 in an interleaved store, as the persistent kernels keep them in LDS)."""
    from kerrray_amd.backends import HostBatch, kernel_params
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import kr_image

    hostref.krh_trace_image_stored.restype = ctypes.c_int
    hostref.krh_trace_image_stored.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-6, atol=1e-8)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=128)
    p = kernel_params(st, integ, term)
    img = kr_image(st, cam.radius, cam.inclination_deg, cam.fov, cam.resolution)
    out = HostBatch(cam.n_rays, np.dtype(np.float64), False)
    assert hostref.krh_trace_image_stored(ctypes.addressof(p), ctypes.addressof(img), ctypes.addressof(out.outputs), 0) == 0
    same(out.result(0.0), run_image(hostref, st, cam, integ, term))


def test_shared_reciprocal_quotients_are_the_ieee_quotients(hostref) -> None:
    """div_by (one reciprocal per denominator, two Markstein corrections) against a / b, bit for bit.

    Random significands over a wide exponent range, signed zeros, powers of
    two, all-ones significands and the operand ranges the metric sees.
    """
    hostref.krh_div_by_mismatches.restype = ctypes.c_int64
    hostref.krh_div_by_mismatches.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    rng = np.random.default_rng(11)
    n = 4_000_000
    cases = []
    m = rng.uniform(1.0, 2.0, (2, n)) * rng.choice([-1.0, 1.0], (2, n))
    e = rng.integers(-300, 300, (2, n))
    cases.append((np.ldexp(m[0], e[0]), np.ldexp(m[1], e[1] // 2)))
    cases.append((rng.uniform(-1e6, 1e6, n), rng.uniform(1e-12, 1e12, n)))
    ones = np.nextafter(2.0, 0.0)  # significand of all ones
    special_b = np.array([1.0, 2.0, 0.5, 3.0, ones, -ones, 1.0 + 2**-52, 7.0, -1.0])
    special_a = np.array([0.0, -0.0, 1.0, -1.0, ones, 3.0, 1.0 + 2**-52, 1e-300, -5e300])
    A, Bm = np.meshgrid(special_a, special_b)
    cases.append((A.ravel(), Bm.ravel()))
    cases.append((rng.uniform(1.0, 2.0, n), np.full(n, ones)))
    for a, b in cases:
        a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
        assert hostref.krh_div_by_mismatches(a.ctypes.data, b.ctypes.data, a.size) == 0


def test_bulk_controller_power_and_clamps(hostref) -> None:
    """The FP32-bulk controller of the precision form: en^(-1/5) within 2 ulp of the extended-precision value on
    [1e-5, 1e5] (the host seeds it from glibc's float log2/exp2, the device from the hardware ones; two Newton
    corrections make either seed converge), and the step factor equal to the exact controller's wherever the
    clamps decide (en outside [1e-5, 1e5], en = 0, en below ~1.9e-4 or above ~4.1) and within 2 ulp elsewhere."""
    hostref.krh_bulk_controller.restype = None
    hostref.krh_bulk_controller.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    rng = np.random.default_rng(12)
    x = np.concatenate([np.logspace(-5, 5, 200_001), 10.0 ** rng.uniform(-5, 5, 200_000),
                        np.array([1e-5, 1e5, 1.0, 2.0, 0.5, np.nextafter(1.0, 0), np.nextafter(1e-5, 1)]),
                        np.array([0.0, 1e-300, 1e-9, np.nextafter(1e-5, 0), np.nextafter(1e5, 2e5), 1e9, 1e300])])
    x = np.ascontiguousarray(x)
    out = np.empty((x.size, 3))
    hostref.krh_bulk_controller(x.ctypes.data, x.size, out.ctypes.data)
    inside = (x >= 1e-5) & (x <= 1e5)
    exact_pow = np.asarray(x[inside], dtype=np.longdouble) ** np.longdouble(-0.2)
    rel = np.abs((out[inside, 0] - exact_pow) / exact_pow).astype(np.float64)
    assert rel.max() <= 2 * 2.0**-52, rel.max()
    exact_f, bulk_f = out[:, 1], out[:, 2]
    clamped = (exact_f == 5.0) | (exact_f == 0.2)
    assert np.array_equal(bulk_f[~inside], exact_f[~inside])
    assert np.array_equal(bulk_f[clamped], exact_f[clamped])
    assert np.max(np.abs(bulk_f - exact_f) / exact_f) <= 2 * 2.0**-52
