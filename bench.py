"""Throughput of the per-photon null-geodesic path on one MI355X.

Workload (one "step"): one 512 x 512 Bardeen image, a = 0.9, i = 60 deg,
r_o = 1000 M, alpha, beta in [-10, 10] M, Dormand-Prince 5(4) at rtol 1e-6 /
atol 1e-8 with the reference's controller, step cap and termination rules,
FP64. Initial conditions are generated on the device; inputs (the image
description) and outputs (states, codes, counters: ~120 B per photon) live
in HBM for the whole timed region. ``value`` = photons per second.

Also reported: the kernel's average duration from HIP events on the stream it
runs on, the algorithmic FP64 rate against the MI355X vector peak, whether
the mask equals the reference's (tests/golden), and the CPU oracle's rate
on this host's cores (a reported baseline, not the target).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s per GPU at 512² a=0.9 rtol1e-6: FP64 vs FP32+FP64-repair, mask-identical"
FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X published FP64 vector peak
SPIN, INCL, R_O, FOV, RES, RTOL, ATOL = 0.9, 60.0, 1000.0, 10.0, 512, 1e-6, 1e-8


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_outputs(torch, n: int, dtype):
    from kerrray_amd import _lib

    dev = torch.device("cuda:0")
    bufs = {
        "Y": torch.empty((n, 8), dtype=dtype, device=dev),
        "lam": torch.empty(n, dtype=torch.float64, device=dev),
        "n_steps": torch.empty(n, dtype=torch.int64, device=dev),
        "n_rejected": torch.empty(n, dtype=torch.int64, device=dev),
        "state": torch.empty(n, dtype=torch.int64, device=dev),
        "m0": torch.empty(n, dtype=torch.float64, device=dev),
        "m1": torch.empty(n, dtype=torch.float64, device=dev),
        "m2": torch.empty(n, dtype=torch.float64, device=dev),
        "m3": torch.empty(n, dtype=torch.float64, device=dev),
    }
    o = _lib.KrOutputs()
    o.Y = bufs["Y"].data_ptr()
    o.lam = bufs["lam"].data_ptr()
    o.n_steps = bufs["n_steps"].data_ptr()
    o.n_rejected = bufs["n_rejected"].data_ptr()
    o.state = bufs["state"].data_ptr()
    o.max_null_error = bufs["m0"].data_ptr()
    o.max_energy_drift = bufs["m1"].data_ptr()
    o.max_lz_drift = bufs["m2"].data_ptr()
    o.max_carter_drift = bufs["m3"].data_ptr()
    return bufs, o


def golden_mask() -> np.ndarray | None:
    path = os.path.join(ROOT, "tests", "golden", "bardeen512_a09_i60.npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        return z["numpy_state"].astype(np.int64)


def cpu_baseline(n_threads: int) -> dict:
    """The CPU oracle (C++ restatement of the reference kernel) on this host: full image, best of 3."""
    from oracle import oracle as O
    from types import SimpleNamespace as NS

    O.build()
    st = NS(mass=1.0, spin=SPIN)
    integ = NS(method="rk45", rtol=RTOL, atol=ATOL, step_size=0.01, max_steps=100_000, lambda_max=1e4,
               h_min=1e-12, dtype="float64", h_max=None)
    term = NS(horizon_epsilon=1e-6, escape_radius=1000.0)
    Y0 = O.bardeen_image(1.0, SPIN, R_O, INCL, FOV, RES)
    best = math.inf
    for _ in range(3):
        t0 = time.perf_counter()
        res = O.integrate_batch(st, Y0, integ, term, n_threads=n_threads, dynamic=True)
        best = min(best, time.perf_counter() - t0)
    ref = golden_mask()
    mask_ok = None if ref is None else bool(np.array_equal(np.asarray(res.state, np.int64), ref))
    static = math.inf
    for _ in range(2):
        t0 = time.perf_counter()
        O.integrate_batch(st, Y0, integ, term, n_threads=n_threads, dynamic=False)
        static = min(static, time.perf_counter() - t0)
    one = Y0[::64]
    t1 = math.inf
    for _ in range(2):
        t0 = time.perf_counter()
        O.integrate_batch(st, one, integ, term, n_threads=1)
        t1 = min(t1, time.perf_counter() - t0)
    return {
        "value": len(Y0) / best,
        "unit": "photons/s",
        "cores": n_threads,
        "kind": "port",
        "label": "Numba-equivalent restatement (oracle/kr_oracle.cpp, the reference's numba_kernel.py in C++, "
                 "FP64, -O2 -fno-fast-math -ffp-contract=off)",
        "mask_identical_to_reference": mask_ok,
        "sample": f"the full 512x512 workload (262144 photons), best of 3, OpenMP dynamic over photons on {n_threads} threads",
        "static_schedule": {"value": len(Y0) / static,
                            "sample": f"the same workload, best of 2, OpenMP static over photons on {n_threads} threads"},
        "single_core": {"value": len(one) / t1, "sample": "every 64th pixel (4096 photons), 1 thread, best of 2"},
        "host": host_cpu(),
    }


EXP009_RAYS, EXP009_CPU_SAMPLE = 1_000_000, 100_000


def exp009_points(n: int) -> np.ndarray:
    """The seeded image-plane points of the reference's CPU throughput benchmark (benchmarks/cpu.py:157-170):
    uniform in [-12, 12]^2 M, drawn as (n, 2) pairs from a PCG64 generator with seed 0."""
    return np.random.default_rng(0).uniform(-12.0, 12.0, size=(n, 2))


def exp009_options():
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions

    return IntegratorOptions(method="rk45", rtol=1e-9, atol=1e-11), TerminationOptions()


def exp009_report(hip, log, n_threads: int | None) -> dict:
    """The reference's one published throughput workload (EXP-009, docs/numerical_analysis.md:266-301): 1e6
    seeded rays, ZAMO camera at r_o = 1000 M, i = 60 deg, a = 0.9, rk45 rtol 1e-9 / atol 1e-11.  On hip the
    states are built on the device (kr_point_states_dev) and traced through kr_integrate_batch_dev, the
    path the reference's benchmark takes through Backend.integrate_batch; the CPU leg runs the oracle on the
    first rays of the same sequence."""
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.camera import Camera, image_plane_states_device
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    cam = Camera(radius=R_O, inclination_deg=60.0, fov=12.0, resolution=2)
    integ, term = exp009_options()
    pts = exp009_points(EXP009_RAYS)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    y0 = image_plane_states_device(cam, st, pts[:, 0], pts[:, 1], backend=hip)
    e1.record(stream)
    bufs, outs = device_outputs(torch, EXP009_RAYS, torch.float64)
    p = kernel_params(st, integ, term)

    def trace():
        _lib.check(_lib.lib().kr_integrate_batch_dev(hip.handle.ptr, ctypes.byref(p), y0.data_ptr(), EXP009_RAYS,
                                                     ctypes.byref(outs), ctypes.c_void_p(stream.cuda_stream)),
                   "kr_integrate_batch_dev")

    trace()
    ms = timed(trace, 2)
    torch.cuda.synchronize()
    states_ms = e0.elapsed_time(e1)
    state = bufs["state"].cpu().numpy()
    counts = np.bincount(state, minlength=7)
    out = {"workload": "EXP-009: 1e6 rays, seeded uniform alpha,beta in [-12,12]^2 M (PCG64 seed 0), ZAMO camera "
                       "r_o=1000 M, i=60deg, a=0.9, rk45 rtol 1e-9 atol 1e-11, FP64",
           "reference_published": {"rays_per_s": 7949, "source": "docs/numerical_analysis.md:292-301 (Numba, 8 "
                                   "logical CPUs of a laptop, N = 1e6)"},
           "hip": {"rays": EXP009_RAYS, "trace_kernel_ms": ms, "states_kernel_ms": states_ms,
                   "rays_per_s": EXP009_RAYS / (ms * 1e-3), "captured": int(counts[2]), "escaped": int(counts[1]),
                   "other": int(EXP009_RAYS - counts[1] - counts[2]),
                   "mean_steps": float(bufs["n_steps"].double().mean().item())}}
    log(f"EXP-009 1e6 rays at rtol 1e-9: {ms:.1f} ms ({out['hip']['rays_per_s']:.3g} rays/s)")
    del bufs, y0
    torch.cuda.empty_cache()
    if n_threads:
        from oracle import oracle as O

        sample = pts[:EXP009_CPU_SAMPLE]
        Y0, bad = O.zamo_states(1.0, 0.9, R_O, cam.inclination_rad, 0.0, sample[:, 0], sample[:, 1])
        t0 = time.perf_counter()
        res = O.integrate_batch(st, Y0, integ, term, n_threads=n_threads, dynamic=True)
        sec = time.perf_counter() - t0
        same = bool(np.array_equal(np.asarray(res.state, np.int64), state[:EXP009_CPU_SAMPLE]))
        out["cpu_baseline"] = {"rays_per_s": len(Y0) / sec, "cores": n_threads, "kind": "port",
                               "sample": f"the first {EXP009_CPU_SAMPLE} rays of the same sequence, one run",
                               "classes_identical_to_hip": same}
        log(f"EXP-009 CPU oracle: {len(Y0) / sec:.3g} rays/s on {n_threads} threads")
    return out


def host_cpu() -> dict:
    """The host's CPU model (from /proc/cpuinfo), its logical CPU count and this process's CPU affinity."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for row in f:
                if row.startswith("model name"):
                    model = row.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def persistent_single(hip, ref_state) -> dict:
    """The 512^2 image through the persistent work-stealing kernel (one job, hipGraph replay)."""
    from kerrray_amd import flops
    from kerrray_amd.batch import DeviceBatch, ImageJob

    batch = DeviceBatch(hip, [ImageJob(SPIN, INCL, RES, RTOL, ATOL, fov=FOV, r_o=R_O)])
    batch.run(graph=True)
    (res,) = batch.results()
    ms = batch.timed(repeats=10, graph=True)
    batch.close()
    fl = flops.batch_flops(res.n_steps, res.n_rejected)
    ach = fl / (ms * 1e-3) / 1e12
    return {"photons_per_s": RES * RES / (ms * 1e-3), "kernel_ms": ms, "roofline_frac": ach / FP64_VECTOR_PEAK_TFLOPS,
            "mask_identical_to_reference": None if ref_state is None else bool(np.array_equal(res.state, ref_state)),
            "tune": {"block_size": 256, "waves_per_simd": 2, "refill_interval": 4}}


class PlannedRunner:
    """One planned-trace form of the 512^2 image on device buffers (launch stream = torch's current stream)."""

    def __init__(self, hip, fp32_bulk: bool, st, integ, term, res: int = RES, spin: float = SPIN, incl: float = INCL,
                 camera=None):
        import torch

        from kerrray_amd import _lib
        from kerrray_amd.backends import kernel_params
        from kerrray_amd.boundary import finite_distance_scale
        from kerrray_amd.camera import image_struct
        from kerrray_amd.planned import EdgeTable, PlannedOptions, edge_table
        from kerrray_amd.shadow import kr_image

        self.hip, self.n = hip, res * res
        self.opts = PlannedOptions(fp32_bulk=fp32_bulk)
        self.table = edge_table(st, incl)
        img = kr_image(st, R_O, incl, FOV, res)
        if camera is not None:  # a ZAMO camera sees Bardeen's curve shrunk by the finite-distance factor
            k = finite_distance_scale(st, camera.radius)
            self.table = EdgeTable(centre=(self.table.centre[0] / k, self.table.centre[1] / k),
                                   radii=self.table.radii / k)
            img = image_struct(st, camera)
        self.bufs, self.outs = device_outputs(torch, self.n, torch.float64)
        self.radii = torch.from_numpy(self.table.radii).cuda()
        self.nbytes = _lib.lib().kr_planned_workspace_bytes(self.n, self.table.n_angles)
        self.work = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        self.params, self.img, self.ko = kernel_params(st, integ, term), img, self.opts.kr()
        self.stream = torch.cuda.current_stream()

    def __call__(self) -> None:
        from kerrray_amd import _lib

        _lib.check(_lib.lib().kr_trace_image_planned_dev(
            self.hip.handle.ptr, ctypes.byref(self.params), ctypes.byref(self.img), ctypes.byref(self.ko),
            self.table.centre[0], self.table.centre[1], self.radii.data_ptr(), self.table.n_angles,
            ctypes.byref(self.outs), self.work.data_ptr(), self.nbytes, ctypes.c_void_p(self.stream.cuda_stream)),
            "kr_trace_image_planned_dev")

    def stats(self) -> dict:
        """Counters of the last run."""
        from kerrray_amd import _lib

        st = _lib.KrPlannedStats()
        _lib.check(_lib.lib().kr_planned_read_stats(self.work.data_ptr(), self.nbytes, ctypes.byref(st),
                                                    ctypes.c_void_p(self.stream.cuda_stream)), "kr_planned_read_stats")
        _lib.raise_for_flags(st.flags)
        total = int((self.bufs["n_steps"] + self.bufs["n_rejected"]).sum())
        return {"edge_band_pixels": st.planned_band, "repaired_pixels": st.repaired,
                "repair_rounds": st.repair_rounds, "fp64_attempt_fraction": st.fp64_attempts / max(1, total)}

    def options(self) -> dict:
        o = self.opts
        return {"band_px": o.band_px, "horizon_margin": o.horizon_margin, "axis_sin": o.axis_sin,
                "turn_rmax": o.turn_rmax, "turn_drdl": o.turn_drdl, "refill_interval": o.refill_interval,
                "block_size": o.block_size, "longest_first": o.longest_first, "split_roles": o.split_roles,
                "waves_per_simd": o.waves_per_simd or "auto", "grid_fraction": o.grid_fraction or "auto"}


class PlainRunner:
    """The plain one-photon-per-lane FP64 kernel (kr_trace_image_dev)."""

    def __init__(self, hip, st, integ, term):
        import torch

        from kerrray_amd.backends import kernel_params
        from kerrray_amd.shadow import kr_image

        self.hip, self.n = hip, RES * RES
        self.bufs, self.outs = device_outputs(torch, self.n, torch.float64)
        self.params, self.img = kernel_params(st, integ, term), kr_image(st, R_O, INCL, FOV, RES)
        self.stream = torch.cuda.current_stream()

    def __call__(self) -> None:
        from kerrray_amd import _lib

        _lib.check(_lib.lib().kr_trace_image_dev(self.hip.handle.ptr, ctypes.byref(self.params), ctypes.byref(self.img),
                                                 ctypes.byref(self.outs), None, ctypes.c_void_p(self.stream.cuda_stream)),
                   "kr_trace_image_dev")


class RowsRunner:
    """The reference's own path: given initial states (resident in HBM) through kr_integrate_batch_dev, which
    runs FP64 Dormand-Prince batches on the longest-first persistent kernel."""

    def __init__(self, hip, st, integ, term):
        import torch

        from kerrray_amd import _lib
        from kerrray_amd.backends import kernel_params
        from kerrray_amd.shadow import kr_image

        self.hip, self.n = hip, RES * RES
        self.bufs, self.outs = device_outputs(torch, self.n, torch.float64)
        self.y0 = torch.empty((self.n, 8), dtype=torch.float64, device="cuda")
        self.stream = torch.cuda.current_stream()
        img = kr_image(st, R_O, INCL, FOV, RES)
        flags = torch.zeros(1, dtype=torch.int32, device="cuda")
        _lib.check(_lib.lib().kr_image_states_dev(hip.handle.ptr, ctypes.byref(img), 1, self.y0.data_ptr(),
                                                  flags.data_ptr(), ctypes.c_void_p(self.stream.cuda_stream)),
                   "kr_image_states_dev")
        _lib.raise_for_flags(int(flags.item()))
        self.params = kernel_params(st, integ, term)

    def __call__(self) -> None:
        from kerrray_amd import _lib

        _lib.check(_lib.lib().kr_integrate_batch_dev(self.hip.handle.ptr, ctypes.byref(self.params),
                                                     self.y0.data_ptr(), self.n, ctypes.byref(self.outs),
                                                     ctypes.c_void_p(self.stream.cuda_stream)),
                   "kr_integrate_batch_dev")


def end_to_end(hip, st, integ, term, ref_state) -> dict:
    """The same 512^2 image through the host-buffer entry points, wall clock per call (PCIe included):
    the FP32 form through kr_trace_image_planned (compute_shadow_planned), and the drop-in row path the
    reference's own callers take, HipBackend.integrate_batch on host initial states (H2D of Y0, D2H of
    the rows). Reported beside `value`, never as it."""
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned

    cam = BardeenCamera(radius=R_O, inclination_deg=INCL, fov=FOV, resolution=RES)
    n = cam.n_rays

    def wall(fn, repeats: int = 5) -> tuple[float, object]:
        out = fn()
        t0 = time.perf_counter()
        for _ in range(repeats):
            out = fn()
        return 1e3 * (time.perf_counter() - t0) / repeats, out

    ms_p, (img, _) = wall(lambda: compute_shadow_planned(st, cam, integ, term, backend=hip,
                                                         options=PlannedOptions(fp32_bulk=True)))
    y0 = bardeen_initial_states(cam, st)
    ms_b, res = wall(lambda: hip.integrate_batch(st, y0, integ, term))
    same_p = None if ref_state is None else bool(np.array_equal(img.result.state, ref_state))
    same_b = None if ref_state is None else bool(np.array_equal(res.state, ref_state))
    return {"workload": "the timed 512^2 image through host buffers, wall clock per call, mean of 5 after one "
                        "warm-up call (fresh host output arrays each call, as the reference's callers get)",
            "fp32_fp64_repair_host_ms": ms_p, "fp32_fp64_repair_host_photons_per_s": n / (ms_p * 1e-3),
            "fp32_fp64_repair_host_mask_identical": same_p,
            "integrate_batch_host_y0_ms": ms_b, "integrate_batch_host_y0_photons_per_s": n / (ms_b * 1e-3),
            "integrate_batch_host_y0_mask_identical": same_b}


def critical_path(hip, st, integ, term, n_steps, n_rej) -> dict:
    """The photon with the most attempts in the timed image, traced alone in exact FP64
    (kr_trace_points_dev: one photon, carried by a lane pair that shares out its trigonometry, as the
    timed form's band wavefronts carry it): an approximate floor for a launch of the image, since the
    timed kernel runs that photon with its own register layout and clock."""
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.shadow import kr_image

    attempts = n_steps + n_rej
    k = int(np.argmax(attempts))
    i, j = divmod(k, RES)
    a, b = BardeenCamera(radius=R_O, inclination_deg=INCL, fov=FOV, resolution=RES).pixel_coordinates()
    alpha = torch.tensor([a[i, j]], dtype=torch.float64, device="cuda")
    beta = torch.tensor([b[i, j]], dtype=torch.float64, device="cuda")
    bufs, outs = device_outputs(torch, 1, torch.float64)
    params, img = kernel_params(st, integ, term), kr_image(st, R_O, INCL, FOV, RES)
    stream = torch.cuda.current_stream()

    def one() -> None:
        _lib.check(_lib.lib().kr_trace_points_dev(hip.handle.ptr, ctypes.byref(params), ctypes.byref(img),
                                                  alpha.data_ptr(), beta.data_ptr(), 1, ctypes.byref(outs), None,
                                                  ctypes.c_void_p(stream.cuda_stream)), "kr_trace_points_dev")

    one()
    ms = timed(one, 3)
    alone = int(bufs["n_steps"].item() + bufs["n_rejected"].item())
    return {"pixel": [i, j], "attempts": int(attempts[k]), "attempts_alone_fp64": alone, "alone_ms": ms,
            "us_per_attempt_alone": 1e3 * ms / max(1, alone), "kernel": "kr_trace_points_dev (exact FP64, one lane pair)"}


def timed(run, repeats: int) -> float:
    """Average ms per call from HIP events on the launch stream (torch's current stream)."""
    import torch

    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(repeats):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / repeats


def form_report(run, ref_state, repeats: int = 5) -> dict:
    from kerrray_amd import flops

    run()
    ms = timed(run, repeats)
    state = run.bufs["state"].cpu().numpy()
    fl = flops.batch_flops(run.bufs["n_steps"].cpu().numpy(), run.bufs["n_rejected"].cpu().numpy())
    out = {"photons_per_s": run.n / (ms * 1e-3), "kernel_ms": ms,
           "roofline_frac": fl / (ms * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS,
           "mask_identical_to_reference": None if ref_state is None else bool(np.array_equal(state, ref_state))}
    if isinstance(run, PlannedRunner):
        out.update(run.stats())
        out["options"] = run.options()
    return out


def precision_report(hip, log) -> dict:
    """1024^2, a=0.99, i=85 deg, rtol 1e-6: FP32 bulk + FP64 repair mask against the FP64 mask, and the
    FP64 720-angle edge bisection at rtol 1e-8 against Bardeen's curve."""
    import torch

    from kerrray_amd.bisection import shadow_edge
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
    from kerrray_amd.geometry import Spacetime

    spin, incl, res = 0.99, 85.0, 1024
    st = Spacetime(1.0, spin)
    integ, term = IntegratorOptions(method="rk45", rtol=1e-6, atol=1e-8), TerminationOptions()
    mixed = PlannedRunner(hip, True, st, integ, term, res=res, spin=spin, incl=incl)
    exact = PlannedRunner(hip, False, st, integ, term, res=res, spin=spin, incl=incl)
    mixed()
    exact()
    ms_mixed, ms_exact = timed(mixed, 3), timed(exact, 3)
    m_state = mixed.bufs["state"].cpu().numpy()
    e_state = exact.bufs["state"].cpu().numpy()
    mask_diff = int(np.count_nonzero((m_state == 2) != (e_state == 2)))
    # Escape directions at r = infinity: the final rows of the two forms end
    # at different radii >= r_esc, so their angles there are not comparable.
    from kerrray_amd.adjoint import directions_at_infinity

    m_th, m_ph = (x.cpu().numpy() for x in directions_at_infinity(hip, st, integ, term, mixed.bufs["Y"], mixed.bufs["state"]))
    e_th, e_ph = (x.cpu().numpy() for x in directions_at_infinity(hip, st, integ, term, exact.bufs["Y"], exact.bufs["state"]))
    esc = (m_state == 1) & (e_state == 1)
    dth, dph = np.abs(m_th - e_th)[esc], np.abs(m_ph - e_ph)[esc]
    # Final rows at r >= r_esc: the two forms' last steps land at different radii, so theta and phi there
    # differ by the path between them as well; photons with the same step count are listed on their own.
    mY, eY = mixed.bufs["Y"].cpu().numpy(), exact.bufs["Y"].cpu().numpy()
    same = esc & (mixed.bufs["n_steps"].cpu().numpy() == exact.bufs["n_steps"].cpu().numpy())
    final_diff = {}
    for name, col in (("r", 1), ("theta", 2), ("phi", 3), ("p_r", 5), ("p_theta", 6)):
        d = np.abs(mY[:, col] - eY[:, col]) / np.maximum(1.0, np.abs(eY[:, col]))
        final_diff[f"{name}_rel_median_escaped"] = float(np.median(d[esc]))
        final_diff[f"{name}_rel_max_escaped"] = float(np.max(d[esc]))
        final_diff[f"{name}_rel_max_same_step_count"] = float(np.max(d[same])) if same.any() else None
    final_diff["escaped_with_same_step_count"] = int(same.sum())
    stats = mixed.stats()
    del mixed, exact
    torch.cuda.empty_cache()
    log(f"precision form 1024^2: mixed {ms_mixed:.2f} ms, fp64 {ms_exact:.2f} ms, mask diff {mask_diff}")
    t0 = time.perf_counter()
    edge, _, _, err = shadow_edge(st, incl, R_O, 1e-8, 720, backend=hip, n_probe=EDGE_PROBES)
    edge_s = time.perf_counter() - t0
    log(f"720-angle FP64 edge at rtol 1e-8: {edge_s:.3f} s ({edge.kernel_ms:.1f} ms of kernels), "
        f"max |r - Bardeen| {err['max']:.2e}")
    return {"workload": "1024x1024 Bardeen image, a=0.99, i=85deg, rtol 1e-6, atol 1e-8",
            "fp32_fp64_repair_ms": ms_mixed, "fp64_planned_ms": ms_exact,
            "photons_per_s_fp32_fp64_repair": res * res / (ms_mixed * 1e-3),
            "photons_per_s_fp64": res * res / (ms_exact * 1e-3),
            "mask_diff_vs_fp64": mask_diff,
            "termination_codes_identical_to_fp64": bool(np.array_equal(m_state, e_state)),
            "final_state_difference_vs_fp64": final_diff,
            "escape_direction_error_vs_fp64": {
                "theta_inf_median": float(np.median(dth)), "theta_inf_p99": float(np.quantile(dth, 0.99)),
                "theta_inf_max": float(np.max(dth)), "phi_inf_median": float(np.median(dph)),
                "phi_inf_p99": float(np.quantile(dph, 0.99)), "phi_inf_max": float(np.max(dph))},
            **stats,
            "edge": {"angles": 720, "rtol": 1e-8, "n_probe": EDGE_PROBES, "seconds": edge_s,
                     "kernel_ms": edge.kernel_ms, "failures": edge.failures,
                     "nonmonotone": edge.nonmonotone, "max_bracket_width": float(edge.width.max()),
                     **{f"vs_bardeen_{k}": float(v) for k, v in err.items()}}}


# Probes per direction and round of the edge bisection: with 63, each of the
# 720 directions is one wavefront and all 6 rounds run in one launch at one
# wavefront per SIMD (58-64 ms of kernels, against 63-65 ms for 127 probes in
# 5 rounds of two launches each, whose 720 workgroups of two wavefronts do not
# fit at one per SIMD).
EDGE_PROBES = 63


def adjoint_report(hip, log) -> dict:
    """Escape-direction gradients of a 256^2 image, a=0.9, i=60 deg, rtol 1e-8, with the Jacobian tape
    (default: photon Jacobians accumulated through per-step Jacobians when the tape is built) and, for
    comparison, the checkpointed reverse sweep (segment 32)."""
    import torch

    from kerrray_amd.adjoint import Tape
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions

    cam = BardeenCamera(radius=R_O, inclination_deg=60.0, fov=FOV, resolution=256)
    a, b = cam.pixel_coordinates()
    alpha = torch.as_tensor(np.ascontiguousarray(a.ravel()), device="cuda")
    beta = torch.as_tensor(np.ascontiguousarray(b.ravel()), device="cuda")
    integ, term = IntegratorOptions(method="rk45", rtol=1e-8, atol=1e-10), TerminationOptions()
    one = torch.ones(alpha.numel(), dtype=torch.float64, device="cuda")
    out = {"workload": "256x256 Bardeen image, a=0.9, i=60deg, rtol 1e-8: d(theta_inf, phi_inf)/d(alpha, beta, spin, theta_o)"}
    grads = {}
    for name, segment in (("jacobian_tape", 0), ("checkpoint_tape_segment32", 32)):
        t0 = time.perf_counter()
        tape = Tape(hip, alpha, beta, 0.9, math.radians(60.0), integ=integ, term=term, segment=segment)
        create_s = time.perf_counter() - t0
        grads[name] = tape.vjp(one, one)
        ms = timed(lambda: tape.vjp(one, one), 3)
        esc = int(tape.escaped.sum())
        kernels_ms = tape.record_ms + tape.jacobian_ms
        out["escaped"] = esc
        out["accepted_steps_on_tape"] = tape.steps
        out[name] = {"tape_create_s": create_s, "record_kernel_ms": tape.record_ms,
                     "jacobian_accumulation_kernels_ms": tape.jacobian_ms, "vjp_ms": ms,
                     "gradients_per_s": esc / (ms * 1e-3),
                     "gradients_per_s_with_tape_kernels": esc / ((ms + kernels_ms) * 1e-3),
                     "consistency_flags": tape.flags()}
        tape.close()
        log(f"adjoint 256^2 ({name}): record {tape.record_ms:.1f} ms, Jacobians {tape.jacobian_ms:.1f} ms, "
            f"vjp {ms:.2f} ms for {esc} photons")
    a, b = grads["jacobian_tape"], grads["checkpoint_tape_segment32"]
    out["max_rel_difference_between_tapes"] = float(((a - b).abs() / b.abs().clamp_min(1e-12 * float(b.abs().max())))
                                                    .max())
    return out


def camera_report(hip, log) -> dict:
    """The reference's own image path: a ZAMO camera at 512^2 (a=0.9, i=60 deg, fov 10, rtol 1e-6) through
    the planned FP64 and FP32+FP64-repair forms; a 512^2 disk render (a=0.9, i=75 deg, fov 25, rtol 1e-6,
    disk r_in = ISCO, r_out = 20); the 360-direction mask edge of the ZAMO image."""
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.boundary import extract_boundary_from_mask
    from kerrray_amd.camera import Camera, image_struct
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.renderer import DiskModel, DiskParams, event_options

    st = Spacetime(1.0, SPIN)
    integ, term = IntegratorOptions(method="rk45", rtol=1e-6, atol=1e-8), TerminationOptions()
    cam = Camera(radius=R_O, inclination_deg=INCL, fov=FOV, resolution=RES)
    out: dict = {"workload": "ZAMO camera (camera.py initial_states) 512x512, a=0.9, i=60deg, fov 10, rtol 1e-6; "
                             "disk render 512x512, a=0.9, i=75deg, fov 25; mask edge 360 directions"}
    masks = {}
    for fp32 in (False, True):
        run = PlannedRunner(hip, fp32, st, integ, term, camera=cam)
        run()
        ms = timed(run, 5)
        masks[fp32] = run.bufs["state"].cpu().numpy() == 2
        out["fp32_fp64_repair" if fp32 else "fp64_planned"] = {
            "kernel_ms": ms, "photons_per_s": run.n / (ms * 1e-3), **run.stats()}
        del run
    out["fp32_fp64_repair"]["mask_identical_to_fp64"] = bool(np.array_equal(masks[False], masks[True]))
    a, b = cam.pixel_coordinates()
    edge = lambda: extract_boundary_from_mask(a, b, masks[False].reshape(RES, RES), n_angles=360, backend=hip)  # noqa: E731
    edge()
    t0 = time.perf_counter()
    _, radii = edge()
    out["mask_edge_360_s"] = time.perf_counter() - t0
    out["mask_edge_mean_radius"] = float(radii.mean())
    # Disk render on device buffers.
    dcam = Camera(radius=R_O, inclination_deg=75.0, fov=25.0, resolution=RES)
    disk = DiskModel.from_params(st, DiskParams())
    params, img, kd = kernel_params(st, integ, term, event_options(disk)), image_struct(st, dcam), disk.kr()
    n = dcam.n_rays
    bufs, outs = device_outputs(torch, n, torch.float64)
    ev_Y = torch.empty((n, 8), dtype=torch.float64, device="cuda")
    ev_hit = torch.empty(n, dtype=torch.uint8, device="cuda")
    outs.event_Y, outs.event_hit = ev_Y.data_ptr(), ev_hit.data_ptr()
    shade = torch.empty((3, n), dtype=torch.float64, device="cuda")
    flags = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    def render() -> None:
        _lib.check(_lib.lib().kr_render_disk_dev(hip.handle.ptr, ctypes.byref(params), ctypes.byref(img),
                                                 ctypes.byref(kd), ctypes.byref(outs), shade[0].data_ptr(),
                                                 shade[1].data_ptr(), shade[2].data_ptr(), flags.data_ptr(),
                                                 ctypes.c_void_p(stream.cuda_stream)), "kr_render_disk_dev")

    render()
    ms = timed(render, 5)
    hit = ev_hit.cpu().numpy().astype(bool)
    g = shade[1].cpu().numpy()[hit]
    out["disk_render"] = {"kernel_ms": ms, "photons_per_s": n / (ms * 1e-3), "disk_hits": int(hit.sum()),
                          "g_min_max": [float(g.min()), float(g.max())], "flags": int(flags.item())}
    log(f"ZAMO 512^2: planned FP64 {out['fp64_planned']['kernel_ms']:.2f} ms, FP32+repair "
        f"{out['fp32_fp64_repair']['kernel_ms']:.2f} ms; disk render {ms:.2f} ms")
    return out


def traffic() -> dict | None:
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def headline_batch(hip, ref_state, log, copies: int = 8) -> dict:
    """The timed image `copies` times over in one batch (hipGraph replay): FP64, and the precision form, whose
    exact photons (the edge band of every copy) and FP32 photons run as two launches side by side.  What a
    caller rendering many frames gets per image; not the timed value, which traces one image per step."""
    from kerrray_amd import _lib
    from kerrray_amd.batch import DeviceBatch, ImageJob
    from kerrray_amd.planned import PlannedOptions

    jobs = [ImageJob(SPIN, INCL, RES, RTOL, ATOL, fov=FOV, r_o=R_O)] * copies
    out: dict = {"workload": f"{copies} copies of the timed 512x512 image in one batch (hipGraph replay), "
                             "tune block 128, 2 waves/SIMD, refill 8"}
    for name, prec in (("fp64", None), ("fp32_fp64_repair", PlannedOptions())):
        b = DeviceBatch(hip, jobs, _lib.KrTune(128, 2, 8, 1), precision=prec)
        b.run(graph=True)
        ms = b.timed(repeats=5, graph=True)
        state = b.bufs["state"].cpu().numpy().reshape(copies, -1)
        same = ref_state is not None and all(np.array_equal(state[k], ref_state) for k in range(copies))
        out[name] = {"kernel_ms": ms, "ms_per_image": ms / copies,
                     "photons_per_s": copies * RES * RES / (ms * 1e-3), "masks_identical_to_reference": bool(same)}
        b.close()
    log(f"{copies} copies of the image in one batch: FP64 {out['fp64']['ms_per_image']:.3f} ms/image, "
        f"FP32+repair {out['fp32_fp64_repair']['ms_per_image']:.3f} ms/image")
    return out


def ragged_batch(hip, log) -> dict:
    """64 ragged images in one persistent launch: autotuned shape, hipGraph replay, checked against the plain kernel."""
    import ctypes as _ct

    from kerrray_amd import _lib, flops
    from kerrray_amd.backends import HostBatch, kernel_params
    from kerrray_amd.batch import DeviceBatch, autotune, ragged_workload
    from kerrray_amd.shadow import kr_image

    jobs = ragged_workload(0, 64)
    t0 = time.perf_counter()
    tune, table = autotune(hip, jobs, log=None)
    tune_s = time.perf_counter() - t0
    log(f"autotune {len(table)} shapes in {tune_s:.1f}s: block={tune.block_size} waves={tune.waves_per_simd} "
        f"refill={tune.refill_interval}")
    batch = DeviceBatch(hip, jobs, tune)
    batch.run(graph=True)
    res = batch.results()
    ms = batch.timed(repeats=5, graph=True)
    fp64_state = batch.bufs["state"].clone()
    batch.close()
    # The precision form on the same batch (FP32 bulk where rtol >= 1e-6, FP64 band, verify/repair).
    from kerrray_amd.planned import PlannedOptions

    mixed = DeviceBatch(hip, jobs, tune, precision=PlannedOptions())
    mixed.run(graph=True)
    ms_mixed = mixed.timed(repeats=5, graph=True)
    pst = mixed.stats()
    same_masks = bool(((mixed.bufs["state"] == 2) == (fp64_state == 2)).all().item())
    mixed.close()
    log(f"precision form on the batch: {ms_mixed:.1f} ms against {ms:.1f} ms FP64, masks identical {same_masks}")
    in_job_order = DeviceBatch(hip, jobs, _lib.KrTune(tune.block_size, tune.waves_per_simd, tune.refill_interval, 0))
    in_job_order.run(graph=True)
    ms_job_order = in_job_order.timed(repeats=3, graph=True)
    in_job_order.close()
    total = sum(j.n_rays for j in jobs)
    fl = sum(flops.batch_flops(r.n_steps, r.n_rejected) for r in res)
    ach = fl / (ms * 1e-3) / 1e12
    # Same images, one plain launch each, for the mask check and the plain-kernel time of the batch.
    plain_ms = 0.0
    identical = True
    for job, r in zip(jobs, res):
        st, integ, term = job.options()
        out = HostBatch(job.n_rays, np.dtype(np.float64), False)
        msv = _ct.c_double(0.0)
        _lib.check(_lib.lib().kr_trace_image(hip.handle.ptr, _ct.byref(kernel_params(st, integ, term)),
                                             _ct.byref(kr_image(st, job.r_o, job.inclination_deg, job.fov,
                                                                job.resolution)),
                                             _ct.byref(out.outputs), _ct.byref(msv)), "kr_trace_image")
        plain_ms += msv.value
        identical &= bool(np.array_equal(out.state, r.state) and np.array_equal(out.Y, r.Y))
    return {"photons": total, "photons_per_s": total / (ms * 1e-3), "kernel_ms": ms,
            "plain_kernel_ms_sum": plain_ms, "speedup_vs_plain_per_image": plain_ms / ms,
            "kernel_ms_images_in_job_order": ms_job_order,
            "roofline_frac": ach / FP64_VECTOR_PEAK_TFLOPS, "masks_identical_to_plain_kernel": identical,
            "tune": {"block_size": tune.block_size, "waves_per_simd": tune.waves_per_simd,
                     "refill_interval": tune.refill_interval, "candidates": len(table), "seconds": tune_s},
            "precision_form": {"kernel_ms": ms_mixed, "photons_per_s": total / (ms_mixed * 1e-3),
                               "masks_identical_to_fp64_batch": same_masks, "fp32_min_rtol": 1e-6,
                               "fp64_band_photons": pst.planned_band, "repaired": pst.repaired,
                               "repair_rounds": pst.repair_rounds, "fp64_attempts": pst.fp64_attempts}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the precision and adjoint reports")
    ap.add_argument("--form", choices=["auto", "fp64_plain", "fp64_planned", "fp32_fp64_repair"], default="auto")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 256 MiB device-to-device once (the reference kernel for the HBM counters)")
    ap.add_argument("--batch-only", metavar="BLOCK,WAVES,REFILL",
                    help="only run the 64-image ragged batch once at this shape (no autotune; for profilers)")
    ap.add_argument("--precision-only", choices=["fp32", "fp64"],
                    help="only trace the 1024^2, a=0.99, i=85 deg image in this planned form, three times (for profilers)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        x = torch.ones(1 << 28, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": x.numel()}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the per-photon path has no cross-GPU step "
              "(photons are independent); run with --gpus 1", file=sys.stderr)
        return 2
    import torch

    from kerrray_amd import flops
    from kerrray_amd.backends import HipBackend
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
    from kerrray_amd.geometry import Spacetime

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    torch.cuda.init()
    log(f"device {torch.cuda.get_device_name(0)}")
    hip = HipBackend(0)
    if args.batch_only:
        from kerrray_amd import _lib
        from kerrray_amd.batch import DeviceBatch, ragged_workload

        shape = [int(v) for v in args.batch_only.split(",")]
        batch = DeviceBatch(hip, ragged_workload(0, 64), _lib.KrTune(*shape, 1))
        ms = batch.timed(repeats=1, graph=False)
        batch.check_flags()
        print(json.dumps({"batch_only": shape, "photons": batch.n, "kernel_ms": ms}), flush=True)
        batch.close()
        return 0
    if args.precision_only:
        runner = PlannedRunner(hip, args.precision_only == "fp32", Spacetime(1.0, 0.99),
                               IntegratorOptions(method="rk45", rtol=1e-6, atol=1e-8), TerminationOptions(),
                               res=1024, spin=0.99, incl=85.0)
        runner()
        ms = timed(runner, 2)
        print(json.dumps({"precision_only": args.precision_only, "photons": runner.n, "kernel_ms": ms}), flush=True)
        return 0
    st = Spacetime(1.0, SPIN)
    integ = IntegratorOptions(method="rk45", rtol=RTOL, atol=ATOL)
    term = TerminationOptions()
    ref = golden_mask()
    n = RES * RES

    runners = {"fp64_plain": PlainRunner(hip, st, integ, term),
               "fp64_rows": RowsRunner(hip, st, integ, term),
               "fp64_planned": PlannedRunner(hip, False, st, integ, term),
               "fp32_fp64_repair": PlannedRunner(hip, True, st, integ, term)}
    forms = {}
    for name, run in runners.items():
        forms[name] = form_report(run, ref)
        log(f"{name}: {forms[name]['kernel_ms']:.3f} ms, mask identical {forms[name]['mask_identical_to_reference']}")
    if args.form == "auto":
        ok = [k for k, v in forms.items() if v["mask_identical_to_reference"] is not False]
        chosen = min(ok, key=lambda k: forms[k]["kernel_ms"])
    else:
        chosen = args.form
    run = runners[chosen]
    log(f"timed form: {chosen}")

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} done")
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        run()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = 1e3 * wall / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, {kernel_ms:.3f} ms/launch (events)")
    if isinstance(run, PlannedRunner):
        run.stats()  # raises on any flag of the timed form (bad initial state, repair limit)
    # Both forms the metric names, each over the same K steps.
    timed_forms = {chosen: {"ms_per_step": ms_per_step, "kernel_ms": kernel_ms,
                            "photons_per_s": n / (ms_per_step * 1e-3)}}
    for other in ("fp64_planned", "fp32_fp64_repair"):
        if other == chosen:
            continue
        orun = runners[other]
        for _ in range(args.warmup):
            orun()
        torch.cuda.synchronize()
        t0o = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            orun()
        ev1.record(stream)
        torch.cuda.synchronize()
        wo = 1e3 * (time.perf_counter() - t0o) / args.steps
        timed_forms[other] = {"ms_per_step": wo, "kernel_ms": ev0.elapsed_time(ev1) / args.steps,
                              "photons_per_s": n / (wo * 1e-3),
                              "mask_identical_to_reference": None if ref is None else
                              bool(np.array_equal(orun.bufs["state"].cpu().numpy(), ref))}
        orun.stats()
        log(f"{other} over the same {args.steps} steps: {wo:.3f} ms/step")
    e2e = end_to_end(hip, st, integ, term, ref)
    log(f"host buffers (PCIe included): FP32 form {e2e['fp32_fp64_repair_host_ms']:.2f} ms per call, "
        f"integrate_batch on host Y0 {e2e['integrate_batch_host_y0_ms']:.2f} ms")
    crit = critical_path(hip, st, integ, term, run.bufs["n_steps"].cpu().numpy(), run.bufs["n_rejected"].cpu().numpy())
    crit["frac_of_ms_per_launch"] = crit["alone_ms"] / kernel_ms
    log(f"longest photon alone: {crit['attempts_alone_fp64']} attempts, {crit['alone_ms']:.3f} ms "
        f"({crit['frac_of_ms_per_launch']:.0%} of a launch)")

    state = run.bufs["state"].cpu().numpy()
    n_steps = run.bufs["n_steps"].cpu().numpy()
    n_rej = run.bufs["n_rejected"].cpu().numpy()
    mask_identical = None if ref is None else bool(np.array_equal(state, ref))
    fl = flops.batch_flops(n_steps, n_rej)
    achieved = fl / (kernel_ms * 1e-3) / 1e12
    counts = np.bincount(state, minlength=7)
    described = {
        "fp64_plain": "FP64, plain one-photon-per-lane kernel",
        "fp64_rows": "FP64, given initial states through kr_integrate_batch_dev (longest-first persistent kernel "
                     "ordered by constants of motion); results bitwise those of the plain kernel",
        "fp64_planned": "FP64, persistent kernel over a longest-first work list (edge band, inside, outside); "
                        "results bitwise those of the plain kernel",
        "fp32_fp64_repair": "FP32 right-hand sides with FP64 state/update/error norm/step control; the edge band "
                            "and polar strip traced in exact FP64 by wavefronts of their own (one per SIMD at "
                            "this size, two lanes per photon sharing its trigonometry); FP64 verify and repair; "
                            "mask identical to the FP64 mask",
    }
    line = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if chosen != "fp32_fp64_repair" else "f32+f64",
        "data": "synthetic: Bardeen image-plane initial conditions generated on the device",
        "config": {
            "workload": "512x512 Bardeen image, a=0.9, i=60deg, r_o=1000M, alpha,beta in [-10,10]M, "
                        "DP5(4) rtol=1e-6 atol=1e-8; " + described[chosen],
            "form": chosen,
            "photons": n,
            "mean_attempts_per_photon": float(np.mean(n_steps + n_rej)),
            "captured": int(counts[2]),
            "escaped": int(counts[1]),
            "other": int(n - counts[1] - counts[2]),
        },
        "mask_identical_to_reference": mask_identical,
        "kernel_ms": kernel_ms,
        "critical_path": crit,
        "roofline": {
            "bound": "valu-fp64",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic_flops_per_launch": fl,
            "launch": {"fp64_plain": "kr_trace_image_dev", "fp64_rows": "kr_integrate_batch_dev (plan + persistent trace)"}
                      .get(chosen, "kr_trace_image_planned_dev (plan + persistent trace + verify + repair kernels)"),
        },
    }
    t = traffic()
    if t is not None and chosen in t:
        line["roofline"]["traffic"] = t[chosen]["bytes_per_launch"]
        line["roofline"]["traffic_source"] = t[chosen]["source"]
    line["forms"] = forms
    line["timed_forms"] = timed_forms
    line["end_to_end"] = e2e
    log("persistent kernel, single image (uniform constants)")
    forms["fp64_persistent"] = persistent_single(hip, ref)
    if not args.no_batch:
        forms["batched_headline_x8"] = headline_batch(hip, ref, log)
        log("ragged batch (64 images, autotuned, hipGraph)")
        forms["batched_ragged64"] = ragged_batch(hip, log)
    if not args.no_extras:
        del runners, run
        torch.cuda.empty_cache()
        line["precision_1024_a099_i85"] = precision_report(hip, log)
        line["adjoint_256_a09_i60"] = adjoint_report(hip, log)
        line["camera_zamo_512_and_disk"] = camera_report(hip, log)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = min(threads, len(os.sched_getaffinity(0)))
    if not args.no_extras:
        line["exp009_context"] = exp009_report(hip, log, None if args.no_cpu_baseline else threads)
    if not args.no_cpu_baseline:
        log(f"cpu baseline on {threads} threads")
        line["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
