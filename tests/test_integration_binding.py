"""The reference-side binding printed in the repository's integration notes, run as it stands.

The module text (``src/kerrray/raytracing/backends/hip_backend.py`` of the
reference, once a maintainer adds it) is taken from the notes' fenced code
block and executed with stand-ins for the four reference modules it imports:

* CPU: ``_kernel_constants`` and ``_tableaus`` hand back the reference's own
  arrays (tests/golden/packing.npz, written by the reference); the binding's
  ``pack_params`` must give every option set the ``kr_params`` bytes that
  ``kerrray_amd.backends.kernel_params`` gives (both are then the reference's
  packing), and its structures must have the header's layout.
* GPU: the binding's ``HipBackend`` must return, bit for bit, what this
  package's backend returns (float64, float32, RK4, the disk event), raise
  ``ValueError`` for ``dop853`` and bad states, and warm up.
"""

from __future__ import annotations

import ctypes
import math
import os
import re
import sys
import types

import numpy as np
import pytest
from conftest import GOLDEN, ROOT, options

from kerrray_amd import _lib
from kerrray_amd.backends import BackendUnavailable, _validate_batch, kernel_params
from kerrray_amd.build import LIB_PATH
from kerrray_amd.geodesics import EventOptions, IntegratorOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.photons import BatchResult


def binding_source() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index("The new module is `src/kerrray/raytracing/backends/hip_backend.py`")
    m = re.compile(r"```python\n(.*?)```", re.S).search(text, start)
    assert m is not None
    return m.group(1)


def load_binding(monkeypatch, kernel_constants, tableaus) -> types.ModuleType:
    """Exec the binding with stand-in reference modules."""
    stubs = {
        "kerrray": types.ModuleType("kerrray"),
        "kerrray.geodesics": types.ModuleType("kerrray.geodesics"),
        "kerrray.photons": types.ModuleType("kerrray.photons"),
        "kerrray.photons.trajectories": types.ModuleType("kerrray.photons.trajectories"),
        "kerrray.raytracing": types.ModuleType("kerrray.raytracing"),
        "kerrray.raytracing.backends": types.ModuleType("kerrray.raytracing.backends"),
        "kerrray.raytracing.backends.base": types.ModuleType("kerrray.raytracing.backends.base"),
        "kerrray.raytracing.backends.numba_backend": types.ModuleType("kerrray.raytracing.backends.numba_backend"),
    }
    stubs["kerrray.geodesics"].IntegratorOptions = IntegratorOptions
    stubs["kerrray.photons.trajectories"].BatchResult = BatchResult
    stubs["kerrray.raytracing.backends.base"].BackendUnavailable = BackendUnavailable
    nb = stubs["kerrray.raytracing.backends.numba_backend"]
    nb._kernel_constants, nb._tableaus, nb._validate_batch = kernel_constants, tableaus, _validate_batch
    for name, mod in stubs.items():
        monkeypatch.setitem(sys.modules, name, mod)
    monkeypatch.setenv("KERRRAY_HIP_LIB", LIB_PATH)
    mod = types.ModuleType("hip_backend")
    exec(compile(binding_source(), "INTEGRATION.md:hip_backend.py", "exec"), mod.__dict__)
    return mod


def constants_from_kernel_params(st, integ, term, events, dtype):
    """This package's packing in the reference's return form (pinned to it by tests/test_packing.py)."""
    p = kernel_params(st, integ, term, events)
    return (np.array(p.geom[:]), np.array(p.consts[:], dtype=dtype), np.array(p.params[:]),
            np.array(p.ipar[:], dtype=np.int64))


def tableaus_from_kernel_params(dtype):
    p = kernel_params(Spacetime(), IntegratorOptions(), options()[1])
    return (np.array(p.tab_a[:]).reshape(7, 7).astype(dtype), np.array(p.tab_e[:], dtype=dtype),
            np.array(p.rk4_b[:], dtype=dtype), np.array(p.rk4_c[:], dtype=dtype))


def test_binding_packs_the_reference_arrays_into_kr_params(monkeypatch) -> None:
    from test_packing import options_of

    with np.load(os.path.join(GOLDEN, "packing.npz"), allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    current = {}

    def ref_constants(st, integ, term, events, dtype):
        assert np.dtype(dtype) == np.float64  # the binding always asks for the float64 packing
        k = current["k"]
        return tuple(gold[f"p{k}_f64_{n}"] for n in ("geom", "consts", "params", "ipar"))

    def ref_tableaus(dtype):
        assert np.dtype(dtype) == np.float64
        k = current["k"]
        return tuple(gold[f"p{k}_f64_{n}"] for n in ("tab_a", "tab_e", "rk4_b", "rk4_c"))

    mod = load_binding(monkeypatch, ref_constants, ref_tableaus)
    assert ctypes.sizeof(mod.KrParams) == ctypes.sizeof(_lib.KrParams)
    assert [f[0] for f in mod.KrParams._fields_] == [f[0] for f in _lib.KrParams._fields_]
    assert [f[0] for f in mod.KrOutputs._fields_] == [f[0] for f in _lib.KrOutputs._fields_]
    for name, _ in mod.KrParams._fields_:
        assert getattr(mod.KrParams, name).offset == getattr(_lib.KrParams, name).offset
    assert mod.ABI_VERSION == _lib.ABI_VERSION
    for k, row in enumerate(gold["sets"]):
        current["k"] = k
        st, integ, term, events = options_of(row)
        got = bytes(mod.pack_params(st, integ, term, events))
        want = bytes(kernel_params(st, integ, term, events))
        assert got == want, k


@pytest.mark.gpu
def test_binding_backend_equals_this_packages_backend(monkeypatch, hip) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states

    mod = load_binding(monkeypatch, constants_from_kernel_params, tableaus_from_kernel_params)
    ref_side = mod.HipBackend(0)
    assert ref_side.name == "hip"
    st = Spacetime(1.0, 0.9)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=75.0, fov=12.0, resolution=24), st)
    cases = [
        ("float64 rk45", options(rtol=1e-7, atol=1e-9), None),
        ("float32 rk45", options(rtol=1e-6, atol=1e-8, dtype="float32", horizon_epsilon=1e-3), None),
        ("rk4 h 0.05", options(method="rk4", step_size=0.05, max_steps=20_000), None),
        ("disk event", options(rtol=1e-8, atol=1e-10), EventOptions(disk_plane=True, r_in=2.32, r_out=20.0)),
        ("float32 disk", options(rtol=1e-6, atol=1e-8, dtype="float32", horizon_epsilon=1e-3),
         EventOptions(disk_plane=True, r_in=6.0, r_out=25.0)),
    ]
    for label, (integ, term), events in cases:
        got = ref_side.integrate_batch(st, Y0, integ, term, events)
        want = hip.integrate_batch(st, Y0, integ, term, events)
        print(f"[binding] {label}: captured {int((got.state == 2).sum())}, disk hits "
              f"{None if got.event_hit is None else int(got.event_hit.sum())}", flush=True)
        for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_energy_drift",
                     "max_lz_drift", "max_carter_drift", "event_Y", "event_hit"):
            a, b = getattr(got, name), getattr(want, name)
            if b is None:
                assert a is None, (label, name)
            else:
                assert a.dtype == b.dtype and np.array_equal(a, b), (label, name)
    integ, term = options()
    with pytest.raises(ValueError, match="dop853"):
        ref_side.integrate_batch(st, Y0, IntegratorOptions(method="dop853"), term)
    with pytest.raises(ValueError, match="finite"):
        ref_side.integrate_batch(st, np.full((2, 8), math.nan), integ, term)
    assert ref_side.warm_up(st, integ, term) >= 0.0
    empty = ref_side.integrate_batch(st, np.zeros((0, 8)), integ, term)
    assert empty.Y.shape == (0, 8)
    ref_side.close()
