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
    path = os.path.join(ROOT, "tests", "golden", "config2_a09_i60_512.npz")
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
        O.integrate_batch(st, Y0, integ, term, n_threads=n_threads, dynamic=True)
        best = min(best, time.perf_counter() - t0)
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
        "sample": f"the full 512x512 workload (262144 photons), best of 3, OpenMP dynamic over photons on {n_threads} threads",
        "single_core": {"value": len(one) / t1, "sample": "every 64th pixel (4096 photons), 1 thread, best of 2"},
    }


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


def planned_image(hip, ref_state, fp32_bulk: bool, repeats: int = 10) -> dict:
    """512^2 image through the planned kernels (device buffers, HIP events on the launch stream)."""
    import ctypes as _ct

    import torch

    from kerrray_amd import _lib, flops
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.planned import PlannedOptions, compute_shadow_planned, edge_table
    from kerrray_amd.shadow import kr_image

    st = Spacetime(1.0, SPIN)
    integ, term = IntegratorOptions(method="rk45", rtol=RTOL, atol=ATOL), TerminationOptions()
    opts = PlannedOptions(fp32_bulk=fp32_bulk)
    cam = BardeenCamera(radius=R_O, inclination_deg=INCL, fov=FOV, resolution=RES)
    image, stats = compute_shadow_planned(st, cam, integ, term, backend=hip, options=opts)
    table = edge_table(st, INCL)
    n = RES * RES
    bufs, outs = device_outputs(torch, n, torch.float64)
    radii = torch.from_numpy(table.radii).cuda()
    nbytes = _lib.lib().kr_planned_workspace_bytes(n, table.n_angles)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    params, img, ko = kernel_params(st, integ, term), kr_image(st, R_O, INCL, FOV, RES), opts.kr()
    stream = torch.cuda.current_stream()

    def step():
        _lib.check(_lib.lib().kr_trace_image_planned_dev(
            hip.handle.ptr, _ct.byref(params), _ct.byref(img), _ct.byref(ko), table.centre[0], table.centre[1],
            radii.data_ptr(), table.n_angles, _ct.byref(outs), work.data_ptr(), nbytes,
            _ct.c_void_p(stream.cuda_stream)), "kr_trace_image_planned_dev")

    step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(repeats):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / repeats
    state = bufs["state"].cpu().numpy()
    n_steps, n_rej = bufs["n_steps"].cpu().numpy(), bufs["n_rejected"].cpu().numpy()
    fl = flops.batch_flops(n_steps, n_rej)
    return {"photons_per_s": n / (ms * 1e-3), "kernel_ms": ms,
            "roofline_frac": fl / (ms * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS,
            "mask_identical_to_reference": None if ref_state is None else bool(np.array_equal(state, ref_state)),
            "edge_band_pixels": stats.planned_band, "repaired_pixels": stats.repaired,
            "fp64_attempt_fraction": stats.fp64_fraction,
            "options": {"band_px": opts.band_px, "horizon_margin": opts.horizon_margin, "axis_sin": opts.axis_sin,
                        "turn_rmax": opts.turn_rmax, "turn_drdl": opts.turn_drdl,
                        "refill_interval": opts.refill_interval, "block_size": opts.block_size}}


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
    batch.close()
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
            "roofline_frac": ach / FP64_VECTOR_PEAK_TFLOPS, "masks_identical_to_plain_kernel": identical,
            "tune": {"block_size": tune.block_size, "waves_per_simd": tune.waves_per_simd,
                     "refill_interval": tune.refill_interval, "candidates": len(table), "seconds": tune_s}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the per-photon path has no cross-GPU step "
              "(photons are independent); run with --gpus 1", file=sys.stderr)
        return 2
    import torch

    from kerrray_amd import _lib, flops
    from kerrray_amd.backends import HipBackend, kernel_params
    from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import kr_image

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    torch.cuda.init()
    log(f"device {torch.cuda.get_device_name(0)}")
    hip = HipBackend(0)
    st = Spacetime(1.0, SPIN)
    integ = IntegratorOptions(method="rk45", rtol=RTOL, atol=ATOL)
    term = TerminationOptions()
    params = kernel_params(st, integ, term)
    img = kr_image(st, R_O, INCL, FOV, RES)
    n = RES * RES
    bufs, outs = device_outputs(torch, n, torch.float64)
    stream = torch.cuda.current_stream()
    L = _lib.lib()

    def step() -> None:
        _lib.check(L.kr_trace_image_dev(hip.handle.ptr, ctypes.byref(params), ctypes.byref(img), ctypes.byref(outs),
                                        None, ctypes.c_void_p(stream.cuda_stream)), "kr_trace_image_dev")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} done")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = 1e3 * wall / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, {kernel_ms:.3f} ms/kernel (events)")

    state = bufs["state"].cpu().numpy()
    n_steps = bufs["n_steps"].cpu().numpy()
    n_rej = bufs["n_rejected"].cpu().numpy()
    ref = golden_mask()
    mask_identical = None if ref is None else bool(np.array_equal(state, ref))
    fl = flops.batch_flops(n_steps, n_rej)
    achieved = fl / (kernel_ms * 1e-3) / 1e12
    counts = np.bincount(state, minlength=7)
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
        "dtype": "f64",
        "data": "synthetic: Bardeen image-plane initial conditions generated on the device",
        "config": {
            "workload": "512x512 Bardeen image, a=0.9, i=60deg, r_o=1000M, alpha,beta in [-10,10]M, "
                        "DP5(4) rtol=1e-6 atol=1e-8, FP64, plain one-photon-per-lane kernel",
            "photons": n,
            "mean_attempts_per_photon": float(np.mean(n_steps + n_rej)),
            "captured": int(counts[2]),
            "escaped": int(counts[1]),
            "other": int(n - counts[1] - counts[2]),
        },
        "mask_identical_to_reference": mask_identical,
        "kernel_ms": kernel_ms,
        "roofline": {
            "bound": "valu-fp64",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic_flops_per_launch": fl,
        },
    }
    log("persistent kernel, single image (uniform constants)")
    line["forms"] = {"fp64_plain": {"photons_per_s": n / (kernel_ms * 1e-3), "kernel_ms": kernel_ms,
                                    "roofline_frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
                                    "mask_identical_to_reference": mask_identical}}
    line["forms"]["fp64_persistent"] = persistent_single(hip, ref)
    log("planned FP64 (longest-first schedule)")
    line["forms"]["fp64_planned"] = planned_image(hip, ref, fp32_bulk=False)
    log("planned FP32 bulk + FP64 repair")
    line["forms"]["fp32_fp64_repair"] = planned_image(hip, ref, fp32_bulk=True)
    if not args.no_batch:
        log("ragged batch (64 images, autotuned, hipGraph)")
        line["forms"]["batched_ragged64"] = ragged_batch(hip, log)
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        threads = min(threads, len(os.sched_getaffinity(0)))
        log(f"cpu baseline on {threads} threads")
        line["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
