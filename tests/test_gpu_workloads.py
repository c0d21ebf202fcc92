"""The two named workloads at their full sizes: the 64-image ragged batch and the 720-angle edge.

* Ragged batch: the 64 images of tests/golden/ragged64_spec.json (PCG64 seed
  0; spins U[0, 0.99], inclinations {17, 60, 85} deg, 64^2 ... 512^2 pixels,
  rtol 1e-5 ... 1e-10 with atol = 1e-2 rtol, 5,828,608 photons), one
  hipGraph launch of the persistent kernel at an autotuned shape. Every
  image's rows must be bit for bit those of the plain per-image kernel.
* Edge: 720 uniform directions at a = 0.99, i = 85 deg, FP64 bisection at
  rtol 1e-8 (the loop of scripts/study_edge_bisection.py:53-68, escaped ->
  hi, else -> lo, with 127 probes per direction and round). No probe may
  fail, every bracket must close to 1e-11 M, and every radius must lie
  within 2e-7 M of Bardeen's curve (about 20 x rtol).
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np
import pytest
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIELDS = ("Y", "lam", "n_steps", "n_rejected", "state", "m0", "m1", "m2", "m3")


def test_ragged64_graph_launch_equals_plain_kernel_per_image(hip) -> None:
    import torch

    from bench import device_outputs
    from kerrray_amd import _lib
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.batch import DeviceBatch, autotune, ragged_workload
    from kerrray_amd.camera import image_struct

    jobs = ragged_workload(0, 64)
    with open(os.path.join(GOLDEN, "ragged64_spec.json")) as f:
        spec = json.load(f)
    assert len(spec["images"]) == 64 and sum(j.n_rays for j in jobs) == 5_828_608
    tune, table = autotune(hip, jobs, block_sizes=(128, 256), waves=(2,), intervals=(4, 8), cache=False)
    print(f"[ragged64] autotuned block {tune.block_size} waves {tune.waves_per_simd} refill {tune.refill_interval}: "
          + ", ".join(f"{r['block_size']}/{r['refill_interval']} {r['ms']:.1f} ms" for r in table), flush=True)
    batch = DeviceBatch(hip, jobs, tune)
    batch.run(graph=True)
    torch.cuda.synchronize()
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    flags = torch.zeros(1, dtype=torch.int32, device="cuda")
    mismatched = []
    for k, job in enumerate(jobs):
        st, integ, term = job.options()
        p, img = kernel_params(st, integ, term), image_struct(st, job.make_camera())
        ref, outs = device_outputs(torch, job.n_rays, torch.float64)
        _lib.check(L.kr_trace_image_dev(hip.handle.ptr, ctypes.byref(p), ctypes.byref(img), ctypes.byref(outs),
                                        flags.data_ptr(), ctypes.c_void_p(stream.cuda_stream)), "kr_trace_image_dev")
        lo, hi = int(batch.offsets[k]), int(batch.offsets[k + 1])
        for name in FIELDS:
            if not torch.equal(batch.bufs[name][lo:hi], ref[name]):
                mismatched.append((k, name))
        del ref
    torch.cuda.synchronize()
    assert int(flags.item()) == 0
    batch.check_flags()
    batch.close()
    assert not mismatched, mismatched[:10]


@pytest.mark.parametrize("n_probe", [63, 127])
def test_edge720_a099_i85_fp64_rtol1e8(hip, n_probe) -> None:
    """63 probes per direction: every round in one launch at one wavefront per SIMD; 127: one launch per round."""
    from kerrray_amd.bisection import shadow_edge
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.99)
    edge, angles, centre, err = shadow_edge(st, 85.0, 1000.0, 1e-8, 720, backend=hip, n_probe=n_probe)
    print(f"[edge720 a=0.99 i=85 n_probe={n_probe}] {edge.kernel_ms:.1f} ms of kernels, failures {edge.failures}, "
          f"nonmonotone "
          f"{edge.nonmonotone}, widest bracket {edge.width.max():.2e}, max |r - Bardeen| {err['max']:.2e}, "
          f"rms {err['rms']:.2e}", flush=True)
    assert angles.size == 720
    assert edge.failures == 0 and edge.nonmonotone == 0
    assert float(edge.width.max()) <= 1e-11
    assert err["max"] < 2e-7
    assert np.all(np.isfinite(edge.radius))


def test_ragged64_precision_form_keeps_every_fp64_mask(hip) -> None:
    """The 64-image workload in the precision form: every image's mask equals the FP64 batch's."""
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch, ragged_workload
    from kerrray_amd.planned import PlannedOptions

    jobs = ragged_workload(0, 64)
    tune = _lib.KrTune(256, 2, 8, 1)
    exact = DeviceBatch(hip, jobs, tune)
    ms_exact = exact.timed(repeats=2)
    ref_state = exact.bufs["state"].clone()
    exact.close()
    mixed = DeviceBatch(hip, jobs, tune, precision=PlannedOptions())
    ms_mixed = mixed.timed(repeats=2)
    st = mixed.stats()
    n = mixed.n
    diff = []
    for k in range(len(jobs)):
        lo, hi = int(mixed.offsets[k]), int(mixed.offsets[k + 1])
        if not torch.equal(mixed.bufs["state"][lo:hi] == 2, ref_state[lo:hi] == 2):
            diff.append(k)
    print(f"[ragged64 precision] {ms_mixed:.1f} ms ({n / (ms_mixed * 1e-3):.3g} photons/s) against FP64 "
          f"{ms_exact:.1f} ms ({n / (ms_exact * 1e-3):.3g}); band {st.planned_band}, repaired {st.repaired} in "
          f"{st.repair_rounds} rounds, fp64 attempts {st.fp64_attempts}", flush=True)
    mixed.close()
    assert not diff, diff
