"""Kernel-constant packing and the float32 working dtype, against vectors the reference wrote.

* tests/golden/packing.npz: the reference's own ``_kernel_constants`` and
  ``_tableaus`` (numba_backend.py:87-128) for seven option sets (float64 and
  float32, rk45 and rk4, h_max, disk events, negative spin), in float64 (the
  contents of the C ABI's ``kr_params``) and in the working dtype (what the
  reference's kernel receives). ``kernel_params`` must reproduce the float64
  arrays bit for bit, and their cast to the working dtype, which is what the
  kernels do with them, must reproduce the working-dtype arrays.
* tests/golden/float32_zamo64.npz: the reference's float32 mode on the image
  of docs/numerical_analysis.md section 7 (ZAMO camera, a = 0.9, i = 60 deg,
  64 x 64, fov 12, rtol 1e-6, horizon_epsilon 1e-3), written by its Numba
  kernel run interpreted and by its NumPy integrator. The oracle's float32
  path and the kernel source built for the host must equal the Numba rows
  bit for bit. The Numba rows reproduce the documented run's classes (558
  captured, 3538 escaped) and escaped-ray figures (null error 4.342e-05,
  Carter drift 6.751e-05) to the printed digits. The documented median
  (1.268e-06) and all-ray maximum (4.585) of the null error were printed by a
  run of a dirty tree (its "git commit a9c864d5 (dirty tree)"); the reference
  as checked out gives 1.2672e-06 and 5.924 through its Numba kernel and
  1.2682e-06 and 3.931 through its NumPy integrator, whose float32 rows differ
  from the kernel's in the last bits (the two paths agree on every class), so
  those two figures pin nothing and are not asserted.
"""

from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import pytest
from conftest import GOLDEN

from kerrray_amd.backends import kernel_params
from kerrray_amd.camera import Camera
from kerrray_amd.geodesics import EventOptions, IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime

NAMES = ("geom", "consts", "params", "ipar", "tab_a", "tab_e", "rk4_b", "rk4_c")


@pytest.fixture(scope="module")
def packing() -> dict:
    with np.load(os.path.join(GOLDEN, "packing.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def options_of(row):
    spin, adaptive, f32, rtol, atol, h0, max_steps, lam_max, h_min, h_max, eps, r_esc, disk, r_in, r_out = row
    integ = IntegratorOptions(method="rk45" if adaptive else "rk4", rtol=rtol, atol=atol, step_size=h0,
                              max_steps=int(max_steps), lambda_max=lam_max, h_min=h_min,
                              h_max=None if math.isnan(h_max) else h_max, dtype="float32" if f32 else "float64")
    term = TerminationOptions(horizon_epsilon=eps, escape_radius=r_esc)
    events = EventOptions(disk_plane=True, r_in=r_in, r_out=r_out) if disk else None
    return Spacetime(1.0, spin), integ, term, events


def packed_arrays(p) -> dict[str, np.ndarray]:
    return {
        "geom": np.array(p.geom[:]), "consts": np.array(p.consts[:]), "params": np.array(p.params[:]),
        "ipar": np.array(p.ipar[:], dtype=np.int64), "tab_a": np.array(p.tab_a[:]).reshape(7, 7),
        "tab_e": np.array(p.tab_e[:]), "rk4_b": np.array(p.rk4_b[:]), "rk4_c": np.array(p.rk4_c[:]),
    }


def test_kernel_params_is_the_reference_packing(packing) -> None:
    sets = packing["sets"]
    assert len(sets) == 7
    for k, row in enumerate(sets):
        st, integ, term, events = options_of(row)
        p = kernel_params(st, integ, term, events)
        got = packed_arrays(p)
        for name in NAMES:
            ref = packing[f"p{k}_f64_{name}"]
            assert got[name].dtype == ref.dtype and np.array_equal(got[name].view(np.uint8), ref.view(np.uint8)), \
                (k, name)
            work = packing[f"p{k}_work_{name}"]
            # geom, params (float64) and ipar keep their type; consts and the tableaus take the working dtype.
            cast = got[name] if name in ("geom", "params", "ipar") else got[name].astype(integ.np_dtype)
            assert cast.dtype == work.dtype and np.array_equal(cast.view(np.uint8), work.view(np.uint8)), (k, name)
        assert p.dtype == (1 if integ.dtype == "float32" else 0)


def test_oracle_constants_are_the_reference_packing(packing, oracle) -> None:
    for k, row in enumerate(packing["sets"]):
        st, integ, term, events = options_of(row)
        geom, consts, params, ipar = oracle.kernel_constants(st, integ, term, events)
        for name, arr in (("geom", geom), ("consts", consts), ("params", params), ("ipar", ipar)):
            ref = packing[f"p{k}_f64_{name}"]
            assert np.array_equal(np.asarray(arr).view(np.uint8), ref.view(np.uint8)), (k, name)


@pytest.fixture(scope="module")
def f32() -> dict:
    with np.load(os.path.join(GOLDEN, "float32_zamo64.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def f32_options():
    return (Spacetime(1.0, 0.9), IntegratorOptions(method="rk45", rtol=1e-6, atol=1e-8, dtype="float32"),
            TerminationOptions(horizon_epsilon=1e-3))


ROWS = ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift", "max_energy_drift",
        "max_lz_drift")


def test_float32_rows_of_the_reference_agree_with_its_documented_run(f32) -> None:
    s = f32["numba_state"]
    esc = s == 1
    assert int((s == 2).sum()) == 558 and int(esc.sum()) == 3538
    m = f32["numba_max_null_error"]
    assert f"{m[esc].max():.4g}" == "4.342e-05"
    assert f"{f32['numba_max_carter_drift'][esc].max():.4g}" == "6.751e-05"
    assert f32["numba_Y"].dtype == np.float32
    # The reference's two CPU paths agree on every class.
    assert np.array_equal(f32["numba_state"], f32["numpy_state"])


def test_zamo_camera_gives_the_reference_initial_states(f32) -> None:
    """The camera kernel's source (built for the host) writes the reference's initial states of this image."""
    import subprocess

    from test_kernel_host_build import HOSTREF

    from kerrray_amd.camera import image_struct

    subprocess.run(["make", "-s", "-C", HOSTREF], check=True, capture_output=True)
    L = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    L.krh_image_states.restype = ctypes.c_int
    L.krh_image_states.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    st, _, _ = f32_options()
    img = image_struct(st, Camera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=64))
    y0 = np.zeros((64 * 64, 8))
    assert L.krh_image_states(ctypes.addressof(img), 1, y0.ctypes.data) == 0
    assert np.array_equal(y0, f32["Y0"])


def test_oracle_float32_path_is_the_reference_float32_kernel(f32, oracle) -> None:
    st, integ, term = f32_options()
    res = oracle.integrate_batch(st, f32["Y0"], integ, term)
    for name in ROWS:
        got, ref = np.asarray(getattr(res, name)), f32[f"numba_{name}"]
        assert got.dtype == ref.dtype and np.array_equal(got, ref), name


def test_kernel_source_float32_path_is_the_reference_float32_kernel(f32) -> None:
    import subprocess

    from test_kernel_host_build import HOSTREF

    subprocess.run(["make", "-s", "-C", HOSTREF], check=True, capture_output=True)
    L = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    L.krh_integrate.restype = ctypes.c_int
    L.krh_integrate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]
    from kerrray_amd.backends import HostBatch

    st, integ, term = f32_options()
    y0 = np.ascontiguousarray(f32["Y0"])
    p = kernel_params(st, integ, term)
    out = HostBatch(len(y0), integ.np_dtype, False)
    assert L.krh_integrate(ctypes.addressof(p), y0.ctypes.data, len(y0), ctypes.addressof(out.outputs), 0) == 0
    res = out.result(0.0)
    for name in ROWS:
        assert np.array_equal(np.asarray(getattr(res, name)), f32[f"numba_{name}"]), name
