"""Write the golden vectors of the per-photon path from the reference itself.

Run in the development container, where the read-only reference checkout is
mounted at /root/reference:

    python tests/golden/make_golden.py --workers 8

What it imports from the reference, and why: the exact discrete trajectory
of the adaptive integrator (accepted/rejected step counts, bitwise final
states, hence the captured mask on pixels next to the shadow edge) is fixed
only by running the reference's own code; closed forms pin the mask only
away from the edge. Two reference entry points are run:

* ``kerrray.geodesics.integrate_batch`` (NumPy batched integrator,
  src/kerrray/geodesics/integrators_batch.py:109-241), the reference's
  numerical reference implementation;
* ``integrate_batch_numba(..., interpreted=True)``
  (src/kerrray/raytracing/backends/numba_backend.py:131-207), the Numba
  kernel source (numba_kernel.py:340-474) executed as plain Python because
  Numba is not installed here. Interpreted, its scalar arithmetic calls the
  same glibc ``sin``/``cos``/``pow`` the compiled kernel calls, so these rows
  are the "Numba" answers.

Initial states come from the reference's Bardeen-constant constructor
``kerrray.benchmarks.rayset._off_equatorial_state`` (rayset.py:158-165) on
the pixel centres of ``Camera.pixel_coordinates`` (camera.py:206-222), with
``sign_theta = -1`` for ``beta >= 0`` and ``+1`` otherwise.

Outputs (compressed .npz, data only) go next to this script.
"""

from __future__ import annotations

import argparse
import hashlib
import math
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def _ref():
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    import kerrray.benchmarks.rayset as rayset
    from kerrray.geodesics import IntegratorOptions, TerminationOptions, integrate_batch
    from kerrray.geometry import Spacetime
    from kerrray.raytracing.backends.numba_backend import integrate_batch_numba
    from kerrray.raytracing.camera import Camera

    return rayset, IntegratorOptions, TerminationOptions, integrate_batch, Spacetime, integrate_batch_numba, Camera


def pixel_grid(fov: float, n: int):
    """alpha, beta (H, W) of the reference camera grid (camera.py:206-222)."""
    *_, Camera = _ref()
    cam = Camera(radius=1000.0, inclination_deg=60.0, fov=fov, resolution=n)
    return cam.pixel_coordinates()


def bardeen_states(spin: float, r_o: float, incl_deg: float, alpha: np.ndarray, beta: np.ndarray) -> np.ndarray:
    rayset, _, _, _, Spacetime, _, _ = _ref()
    st = Spacetime(1.0, spin)
    th = math.radians(incl_deg)
    a, b = alpha.ravel(), beta.ravel()
    out = np.empty((a.size, 8))
    for k in range(a.size):
        out[k] = rayset._off_equatorial_state(st, r_o, th, float(a[k]), float(b[k]), -1 if b[k] >= 0.0 else 1)
    return out


def _numba_interp_chunk(args):
    spin, Y0, rtol, atol = args
    _, IntegratorOptions, TerminationOptions, _, Spacetime, integrate_batch_numba, _ = _ref()
    res = integrate_batch_numba(
        Spacetime(1.0, spin), Y0, IntegratorOptions(method="rk45", rtol=rtol, atol=atol), TerminationOptions(),
        interpreted=True,
    )
    return _fields(res)


def _numpy_chunk(args):
    spin, Y0, rtol, atol = args
    _, IntegratorOptions, TerminationOptions, integrate_batch, Spacetime, _, _ = _ref()
    res = integrate_batch(Spacetime(1.0, spin), Y0, IntegratorOptions(method="rk45", rtol=rtol, atol=atol),
                          TerminationOptions())
    return _fields(res)


def _fields(res) -> dict:
    return {
        "Y": res.Y, "state": res.state, "n_steps": res.n_steps, "n_rejected": res.n_rejected, "lam": res.lam,
        "max_null_error": res.max_null_error, "max_carter_drift": res.max_carter_drift,
    }


def run_parallel(fn, spin, Y0, rtol, atol, workers: int, chunk: int) -> dict:
    jobs = [(spin, Y0[i:i + chunk], rtol, atol) for i in range(0, len(Y0), chunk)]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(fn, jobs))
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def schwarzschild64(workers: int) -> None:
    """Schwarzschild a=0, i=60 deg, r_o=1000, 64x64 on [-10,10]^2, rtol 1e-8 / atol 1e-10: every pixel."""
    spin, incl, r_o, fov, n, rtol, atol = 0.0, 60.0, 1000.0, 10.0, 64, 1e-8, 1e-10
    alpha, beta = pixel_grid(fov, n)
    Y0 = bardeen_states(spin, r_o, incl, alpha, beta)
    t0 = time.time()
    nb = run_parallel(_numba_interp_chunk, spin, Y0, rtol, atol, workers, 64)
    t1 = time.time()
    npy = run_parallel(_numpy_chunk, spin, Y0, rtol, atol, workers, 512)
    t2 = time.time()
    print(f"schwarzschild64: numba-interpreted {t1 - t0:.1f}s, numpy {t2 - t1:.1f}s", flush=True)
    np.savez_compressed(
        os.path.join(HERE, "bardeen64_a0_i60.npz"),
        spin=spin, inclination_deg=incl, r_o=r_o, fov=fov, resolution=n, rtol=rtol, atol=atol,
        Y0=Y0,
        **{f"numba_{k}": v for k, v in nb.items()},
        numpy_state=npy["state"], numpy_n_steps=npy["n_steps"], numpy_n_rejected=npy["n_rejected"], numpy_Y=npy["Y"],
    )


def band_indices(spin: float, incl: float, fov: float, n: int, width_px: float) -> np.ndarray:
    """Row-major pixel indices within ``width_px`` pixels of the reference's Bardeen curve."""
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geometry import Spacetime
    from kerrray.photons.orbits import shadow_curve

    ca, cb = shadow_curve(Spacetime(1.0, spin), incl, 4000)
    alpha, beta = pixel_grid(fov, n)
    d = 2.0 * fov / n
    pts = np.stack([alpha.ravel(), beta.ravel()], axis=1)
    # Distance from each pixel centre to the dense polyline (vertex distance is enough at 4000 points).
    best = np.full(pts.shape[0], np.inf)
    for k in range(0, ca.size, 256):
        dx = pts[:, 0:1] - ca[None, k:k + 256]
        dy = pts[:, 1:2] - cb[None, k:k + 256]
        best = np.minimum(best, np.sqrt(dx * dx + dy * dy).min(axis=1))
    return np.flatnonzero(best <= width_px * d)


def kerr512(workers: int) -> None:
    """Kerr a=0.9, i=60 deg, r_o=1000, 512x512 on [-10,10]^2, rtol 1e-6 / atol 1e-8."""
    spin, incl, r_o, fov, n, rtol, atol = 0.9, 60.0, 1000.0, 10.0, 512, 1e-6, 1e-8
    alpha, beta = pixel_grid(fov, n)
    t0 = time.time()
    Y0 = bardeen_states(spin, r_o, incl, alpha, beta)
    t1 = time.time()
    npy = run_parallel(_numpy_chunk, spin, Y0, rtol, atol, workers, 8192)
    t2 = time.time()
    band = band_indices(spin, incl, fov, n, 2.5)
    sample = np.arange(0, n * n, 251)
    pick = np.union1d(band, sample)
    nb = run_parallel(_numba_interp_chunk, spin, Y0[pick], rtol, atol, workers, 32)
    t3 = time.time()
    print(f"kerr512: IC {t1 - t0:.1f}s, numpy {t2 - t1:.1f}s, numba-interpreted {len(pick)} rays {t3 - t2:.1f}s",
          flush=True)
    state = npy["state"].astype(np.int8)
    np.savez_compressed(
        os.path.join(HERE, "bardeen512_a09_i60.npz"),
        spin=spin, inclination_deg=incl, r_o=r_o, fov=fov, resolution=n, rtol=rtol, atol=atol,
        Y0_sha256=sha(Y0), Y0_pick=Y0[pick], pick=pick, band=band,
        numpy_state=state, numpy_n_steps=npy["n_steps"].astype(np.int32),
        numpy_n_rejected=npy["n_rejected"].astype(np.int32), numpy_Y_pick=npy["Y"][pick],
        **{f"numba_{k}": v for k, v in nb.items()},
    )


def bisection(workers: int) -> None:
    """Edge bisection semantics of scripts/study_edge_bisection.py:53-68 on Bardeen states.

    a=0.9, i=60 deg, r_o=1000, rtol 1e-6 (atol 1e-8), 16 directions (linspace + 0.1 rad) about
    the centroid of the reference's Bardeen curve, bracket [0.98, 1.02] of the curve's radii,
    34 iterations, each iteration one reference NumPy batch; escaped -> hi, else -> lo.
    """
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geometry import Spacetime
    from kerrray.photons.orbits import shadow_curve
    from kerrray.raytracing.boundary import polygon_centroid, polygon_radii

    _, IntegratorOptions, TerminationOptions, integrate_batch, _, _, _ = _ref()
    spin, incl, r_o, rtol, n_dirs, iters = 0.9, 60.0, 1000.0, 1e-6, 16, 34
    st = Spacetime(1.0, spin)
    ca, cb = shadow_curve(st, incl, 40000)
    centre = polygon_centroid(ca, cb)
    angles = np.linspace(0.0, 2 * math.pi, n_dirs, endpoint=False) + 0.1
    r_ana = polygon_radii(ca, cb, centre, angles)
    integ = IntegratorOptions(method="rk45", rtol=rtol, atol=rtol * 1e-2, max_steps=2_000_000, lambda_max=4.0 * r_o)
    term = TerminationOptions(horizon_epsilon=1e-6, escape_radius=r_o)
    lo, hi = 0.98 * r_ana, 1.02 * r_ana
    t0 = time.time()
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        a = centre[0] + mid * np.cos(angles)
        b = centre[1] + mid * np.sin(angles)
        Y0 = bardeen_states(spin, r_o, incl, a, b)
        res = integrate_batch(st, Y0, integ, term)
        escaped = res.state == 1
        hi = np.where(escaped, mid, hi)
        lo = np.where(escaped, lo, mid)
    print(f"bisection: {time.time() - t0:.1f}s", flush=True)
    np.savez_compressed(
        os.path.join(HERE, "bisect_a09_i60_16dirs.npz"),
        spin=spin, inclination_deg=incl, r_o=r_o, rtol=rtol, atol=rtol * 1e-2, max_steps=2_000_000,
        lambda_max=4.0 * r_o, iterations=iters, angles=angles, centre=np.asarray(centre), r_analytic=r_ana,
        lo=lo, hi=hi, curve_alpha=ca[::100], curve_beta=cb[::100],
    )


# ZAMO cameras: (spin, inclination_deg, phi_deg, fov, resolution, r_o).  The
# first two are the 64 x 64 images of docs/shadow.md section 5; the rest cover
# retrograde spin, the clamped axis inclinations, large radii, the equatorial
# plane and an observer inside the ergosphere.
ZAMO_CAMERAS = (
    (0.9, 60.0, 0.0, 12.0, 64, 1000.0),
    (0.0, 60.0, 0.0, 12.0, 64, 1000.0),
    (-0.7, 17.0, 20.0, 12.0, 8, 1000.0),
    (0.998, 85.0, 0.0, 12.0, 8, 1000.0),
    (0.9, 0.0, 0.0, 12.0, 4, 1000.0),
    (0.9, 180.0, 30.0, 12.0, 4, 1000.0),
    (0.5, 150.0, 20.0, 5.5, 10, 1.0e4),
    (0.0, 45.0, 0.0, 12.0, 4, 1.0e6),
    (0.9, 90.0, 0.0, 12.0, 8, 1000.0),
    (0.9, 90.0, 0.0, 1.2, 8, 1.9),
)


def zamo(workers: int) -> None:
    """ZAMO camera states (camera.py:225-381) and the two 64 x 64 shadow images at rtol 1e-9.

    For every camera: the tetrad, metric and inverse metric at the observer
    and ``initial_states`` reversed and physical.  For the first two cameras
    the reference's NumPy integrator traces the reversed states with rk45,
    rtol 1e-9, atol 1e-11 and default termination (the settings of
    docs/shadow.md section 5, whose table gives 558 / 3538 and 608 / 3488
    captured / escaped pixels).
    """
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geometry import Spacetime, inverse_metric, metric
    from kerrray.raytracing.camera import Camera, initial_states, zamo_tetrad

    out: dict[str, np.ndarray] = {"cameras": np.asarray(ZAMO_CAMERAS, dtype=np.float64)}
    t0 = time.time()
    for k, (spin, incl, phi, fov, n, r_o) in enumerate(ZAMO_CAMERAS):
        st = Spacetime(1.0, spin)
        cam = Camera(radius=r_o, inclination_deg=incl, phi_deg=phi, fov=fov, resolution=int(n))
        th = cam.inclination_rad
        out[f"c{k}_theta_o"] = np.float64(th)
        out[f"c{k}_tetrad"] = zamo_tetrad(st, r_o, th)
        out[f"c{k}_metric"] = metric(st, r_o, th)
        out[f"c{k}_inverse_metric"] = inverse_metric(st, r_o, th)
        out[f"c{k}_Y0"] = initial_states(cam, st)
        out[f"c{k}_Y0_physical"] = initial_states(cam, st, reverse=False)
    t1 = time.time()
    for k in (0, 1):
        spin = ZAMO_CAMERAS[k][0]
        res = run_parallel(_numpy_chunk, spin, out[f"c{k}_Y0"], 1e-9, 1e-11, workers, 256)
        out[f"c{k}_numpy_state"] = res["state"].astype(np.int8)
        out[f"c{k}_numpy_n_steps"] = res["n_steps"].astype(np.int32)
        out[f"c{k}_numpy_n_rejected"] = res["n_rejected"].astype(np.int32)
    print(f"zamo: states {t1 - t0:.1f}s, two 64x64 traces at rtol 1e-9 {time.time() - t1:.1f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, "zamo_cameras.npz"), **out)


def _numba_disk_chunk(args):
    spin, Y0, rtol, atol, r_in, r_out = args
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geodesics import EventOptions

    _, IntegratorOptions, TerminationOptions, _, Spacetime, integrate_batch_numba, _ = _ref()
    res = integrate_batch_numba(
        Spacetime(1.0, spin), Y0, IntegratorOptions(method="rk45", rtol=rtol, atol=atol), TerminationOptions(),
        events=EventOptions(disk_plane=True, r_in=r_in, r_out=r_out), interpreted=True,
    )
    out = _fields(res)
    out["event_Y"], out["event_hit"] = res.event_Y, res.event_hit
    return out


def disk(workers: int) -> None:
    """render_disk (renderer.py:163-216) at the measured run of docs/rendering.md section 6.

    a = 0.9, i = 75 deg, 64 x 64, fov 25, r_o 1000, rk45 rtol 1e-9 / atol 1e-11,
    disk defaults (r_in = prograde ISCO, r_out 20, p 3, g4): the table there
    gives 3182 / 54 / 860 ESCAPED / CAPTURED / DISK_HIT pixels.  The reference's
    NumPy backend traces the rays, and its Numba kernel, interpreted, traces
    them again (glibc pow: the bitwise answer for the compiled kernel); the crossing states, g, I_obs and r_hit of
    every pixel are kept, and g and I_obs of crossings outside the retrograde
    photon orbit for a retrograde g3 disk, and g of crossings in a
    Schwarzschild frame.
    """
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geometry import Spacetime
    from kerrray.physics.accretion import DiskModel, DiskParams, observed_intensity
    from kerrray.physics.redshift import redshift_factor
    from kerrray.photons.orbits import isco_radius
    from kerrray.raytracing.camera import Camera, initial_states
    from kerrray.raytracing.renderer import render_disk

    _, IntegratorOptions, TerminationOptions, _, _, _, _ = _ref()
    spin, incl, n, fov = 0.9, 75.0, 64, 25.0
    st = Spacetime(1.0, spin)
    model = DiskModel.from_params(st, DiskParams())
    cam = Camera(radius=1000.0, inclination_deg=incl, fov=fov, resolution=n)
    t0 = time.time()
    img = render_disk(st, cam, IntegratorOptions(method="rk45", rtol=1e-9, atol=1e-11), TerminationOptions(), model,
                      backend="numpy")
    hit = img.hit.ravel()
    y_hit = img.hit_states.reshape(-1, 8)[hit]
    retro = DiskModel(r_in=model.r_in, r_out=model.r_out, emissivity_index=2.5, intensity_law="g3", prograde=False)
    y_retro = y_hit[y_hit[:, 1] > 4.5]  # retrograde gas orbits exist only outside r_ph ~ 3.9 M
    g_retro = redshift_factor(st, y_retro, prograde=False)
    y_schw = y_hit[y_hit[:, 1] > 6.5]
    g_schw = redshift_factor(Spacetime(1.0, 0.0), y_schw)
    t1 = time.time()
    Y0 = initial_states(cam, st)
    jobs = [(spin, Y0[i:i + 64], 1e-9, 1e-11, model.r_in, model.r_out) for i in range(0, len(Y0), 64)]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(_numba_disk_chunk, jobs))
    nb = {k: np.concatenate([p_[k] for p_ in parts]) for k in parts[0]}
    print(f"disk: numpy render {t1 - t0:.1f}s, numba-interpreted trace {time.time() - t1:.1f}s, "
          f"counts {img.counts()}", flush=True)
    np.savez_compressed(
        os.path.join(HERE, "disk_a09_i75_64.npz"),
        spin=spin, inclination_deg=incl, resolution=n, fov=fov, r_o=1000.0, rtol=1e-9, atol=1e-11,
        r_in=model.r_in, r_out=model.r_out, emissivity_index=model.emissivity_index,
        redshift_power=model.redshift_power,
        isco=np.array([isco_radius(Spacetime(1.0, a), prograde=p) for a in (0.0, 0.5, 0.9, -0.7) for p in (True, False)]),
        state=img.state.ravel().astype(np.int8), hit=hit, y_hit=y_hit,
        g=img.g.ravel(), intensity=img.intensity.ravel(), r_hit=img.r_hit.ravel(),
        n_steps=img.result.n_steps.astype(np.int32),
        y_retro=y_retro, g_retro=g_retro, intensity_retro=observed_intensity(retro, y_retro[:, 1], g_retro),
        y_schw=y_schw, g_schw=g_schw,
        numba_state=nb["state"].astype(np.int8), numba_n_steps=nb["n_steps"].astype(np.int32),
        numba_n_rejected=nb["n_rejected"].astype(np.int32), numba_event_hit=nb["event_hit"],
        numba_event_Y_hit=nb["event_Y"][nb["event_hit"].astype(bool)],
    )


def edges(workers: int) -> None:
    """extract_boundary_from_mask (boundary.py:116-170) on reference masks.

    The 512 x 512 Kerr mask (a = 0.9, i = 60 deg, fov 10) and the two 64 x 64 ZAMO masks of
    docs/shadow.md section 5 (fov 12) about their captured centroids, 360
    directions; a 128 x 128 synthetic disk about a given centre, 90 directions.
    """
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.raytracing.boundary import extract_boundary_from_mask, mask_centroid

    with np.load(os.path.join(HERE, "bardeen512_a09_i60.npz")) as z:
        mask2 = z["numpy_state"].reshape(512, 512) == 2
    with np.load(os.path.join(HERE, "zamo_cameras.npz")) as z:
        zamo = [z[f"c{k}_numpy_state"].reshape(64, 64) == 2 for k in (0, 1)]
    out = {}
    cases = [("c2", 10.0, mask2, None, 360), ("zamo0", 12.0, zamo[0], None, 360), ("zamo1", 12.0, zamo[1], None, 360)]
    a128, b128 = pixel_grid(10.0, 128)
    cases.append(("disk", 10.0, (a128 - 0.5) ** 2 + b128**2 < 16.0, (0.5, 0.0), 90))
    for name, fov, mask, centre, n_angles in cases:
        a, b = pixel_grid(fov, mask.shape[0])
        ang, rad = extract_boundary_from_mask(a, b, mask, n_angles=n_angles, centre=centre)
        out[f"{name}_mask"] = np.packbits(mask)
        out[f"{name}_fov"] = np.float64(fov)
        out[f"{name}_n"] = np.int64(mask.shape[0])
        out[f"{name}_centre"] = np.asarray(mask_centroid(a, b, mask) if centre is None else centre, dtype=np.float64)
        out[f"{name}_angles"] = ang
        out[f"{name}_radii"] = rad
    np.savez_compressed(os.path.join(HERE, "mask_edges.npz"), **out)
    print("edges: written", flush=True)


# Critical-impact searches (validation/kerr_bisection.py): spin, r0, rtol,
# bracket, tol, n_probe, directions.
IMPACT_SEARCHES = (
    (0.0, 50.0, 1e-8, 4.0, 7.0, 1e-7, 8, (True,)),
    (0.9, 50.0, 1e-8, 2.0, 8.0, 1e-6, 8, (True, False)),
    (-0.5, 1000.0, 1e-9, 2.0, 9.0, 1e-6, 5, (True, False)),
    (0.998, 50.0, 1e-8, 1.5, 8.0, 1e-5, 1, (True,)),
)


def impacts(workers: int) -> None:
    """photon_from_constants (initial_conditions.py:83-112) on equatorial inward rays, and
    bisect_critical_impacts (kerr_bisection.py:177-300) on the searches above (atol = 1e-2 rtol,
    escape radius = r0)."""
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geodesics import IntegratorOptions, TerminationOptions, photon_from_constants
    from kerrray.geometry import Spacetime
    from kerrray.validation.kerr_bisection import bisect_critical_impacts

    out: dict[str, np.ndarray] = {}
    rows, states = [], []
    for spin in (0.0, 0.9, -0.5, 0.998):
        st = Spacetime(1.0, spin)
        for r0 in (50.0, 1000.0):
            for b in (1.5, 3.1, 5.196152422706632, 7.25):
                for lz in (b, -b):
                    rows.append((spin, r0, lz))
                    states.append(photon_from_constants(st, r0, 0.5 * math.pi, 1.0, lz, 0.0, sign_r=-1, sign_theta=1))
    out["constants_rows"] = np.asarray(rows)
    out["constants_Y0"] = np.asarray(states)
    t0 = time.time()
    for k, (spin, r0, rtol, lo, hi, tol, n_probe, dirs) in enumerate(IMPACT_SEARCHES):
        res = bisect_critical_impacts(
            Spacetime(1.0, spin), IntegratorOptions(method="rk45", rtol=rtol, atol=1e-2 * rtol),
            TerminationOptions(escape_radius=r0), r0=r0, b_lo=lo, b_hi=hi, tol=tol, max_iterations=60,
            n_probe=n_probe, directions=dirs)
        out[f"s{k}_brackets"] = np.asarray([[r.b_lo, r.b_hi, r.iterations, r.n_rays, r.max_steps] for r in res])
    print(f"impacts: {time.time() - t0:.1f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, "critical_impacts.npz"), **out)


def zamo_rows(workers: int) -> None:
    """Final rows of the reference's NumPy integrator on the two documented 64 x 64 ZAMO images (rtol 1e-9).

    The images of zamo(): a = 0.9 and a = 0 at i = 60 deg, fov 12, r_o =
    1000, rk45 rtol 1e-9 / atol 1e-11.  Set beside the oracle's rows (the
    Numba kernel's, bit for bit), they measure how far the reference's own
    two CPU paths drift apart on these images.
    """
    with np.load(os.path.join(HERE, "zamo_cameras.npz"), allow_pickle=False) as z:
        cams = z["cameras"]
        y0 = {k: z[f"c{k}_Y0"] for k in (0, 1)}
    out = {}
    t0 = time.time()
    for k in (0, 1):
        res = run_parallel(_numpy_chunk, float(cams[k][0]), y0[k], 1e-9, 1e-11, workers, 256)
        out[f"c{k}_numpy_Y"] = res["Y"]
        out[f"c{k}_numpy_n_steps"] = res["n_steps"].astype(np.int32)
        out[f"c{k}_numpy_n_rejected"] = res["n_rejected"].astype(np.int32)
    print(f"zamo rows: {time.time() - t0:.1f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, "zamo64_numpy_rows.npz"), **out)


# Option sets whose packed kernel constants are recorded: (spin, method, dtype,
# rtol, atol, step_size, max_steps, lambda_max, h_min, h_max (nan: None),
# horizon_epsilon, escape_radius, disk (0/1), r_in, r_out).
PACKING_SETS = (
    (0.9, "rk45", "float64", 1e-6, 1e-8, 0.01, 100_000, 1e4, 1e-12, float("nan"), 1e-6, 1000.0, 0, 6.0, 20.0),
    (0.9, "rk45", "float32", 1e-6, 1e-8, 0.01, 100_000, 1e4, 1e-12, float("nan"), 1e-3, 1000.0, 0, 6.0, 20.0),
    (0.5, "rk4", "float64", 1e-9, 1e-11, 0.05, 5_000, 2e3, 1e-12, float("nan"), 1e-6, 500.0, 0, 6.0, 20.0),
    (0.7, "rk45", "float64", 1e-8, 1e-10, 0.02, 20_000, 5e3, 1e-10, 0.5, 1e-6, 1000.0, 0, 6.0, 20.0),
    (0.9, "rk45", "float64", 1e-9, 1e-11, 0.01, 100_000, 1e4, 1e-12, float("nan"), 1e-6, 1000.0, 1, 2.3209, 20.0),
    (0.9, "rk45", "float32", 1e-5, 1e-7, 0.01, 100_000, 1e4, 1e-12, float("nan"), 1e-3, 1000.0, 1, 6.0, 25.0),
    (-0.5, "rk45", "float64", 1e-9, 1e-11, 0.01, 100_000, 1e4, 1e-12, 2.0, 1e-6, 50.0, 0, 6.0, 20.0),
)


def packing(workers: int) -> None:
    """The reference's own packing of the kernel constants (numba_backend.py:87-128) for PACKING_SETS.

    Both in float64 (what the C ABI's kr_params carries) and in the working
    dtype (what the reference's kernel receives).
    """
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geodesics import EventOptions, IntegratorOptions, TerminationOptions
    from kerrray.geometry import Spacetime
    from kerrray.raytracing.backends.numba_backend import _kernel_constants, _tableaus

    out: dict[str, np.ndarray] = {}
    rows = []
    for k, (spin, method, dtype, rtol, atol, h0, max_steps, lam_max, h_min, h_max, eps, r_esc, disk, r_in,
            r_out) in enumerate(PACKING_SETS):
        st = Spacetime(1.0, spin)
        integ = IntegratorOptions(method=method, rtol=rtol, atol=atol, step_size=h0, max_steps=max_steps,
                                  lambda_max=lam_max, h_min=h_min, h_max=None if math.isnan(h_max) else h_max,
                                  dtype=dtype)
        term = TerminationOptions(horizon_epsilon=eps, escape_radius=r_esc)
        events = EventOptions(disk_plane=True, r_in=r_in, r_out=r_out) if disk else None
        for tag, dt in (("f64", np.dtype(np.float64)), ("work", integ.np_dtype)):
            geom, consts, params, ipar = _kernel_constants(st, integ, term, events, dt)
            tab_a, tab_e, rk4_b, rk4_c = _tableaus(dt)
            for name, arr in (("geom", geom), ("consts", consts), ("params", params), ("ipar", ipar),
                              ("tab_a", tab_a), ("tab_e", tab_e), ("rk4_b", rk4_b), ("rk4_c", rk4_c)):
                out[f"p{k}_{tag}_{name}"] = np.asarray(arr)
        rows.append([spin, 1.0 if method == "rk45" else 0.0, 1.0 if dtype == "float32" else 0.0, rtol, atol, h0,
                     max_steps, lam_max, h_min, h_max, eps, r_esc, disk, r_in, r_out])
    out["sets"] = np.asarray(rows, dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "packing.npz"), **out)
    print(f"packing: {len(PACKING_SETS)} option sets", flush=True)


def _float32_chunk(args):
    Y0, interpreted = args
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geodesics import IntegratorOptions, TerminationOptions, integrate_batch
    from kerrray.geometry import Spacetime
    from kerrray.raytracing.backends.numba_backend import integrate_batch_numba

    st = Spacetime(1.0, 0.9)
    integ = IntegratorOptions(method="rk45", rtol=1e-6, atol=1e-8, dtype="float32")
    term = TerminationOptions(horizon_epsilon=1e-3)
    if interpreted:
        res = integrate_batch_numba(st, Y0, integ, term, interpreted=True)
    else:
        res = integrate_batch(st, Y0, integ, term)
    out = _fields(res)
    out["max_energy_drift"], out["max_lz_drift"] = res.max_energy_drift, res.max_lz_drift
    return out


def float32_zamo(workers: int) -> None:
    """The reference's float32 mode on the image of docs/numerical_analysis.md section 7 (EXP-011).

    ZAMO camera, a = 0.9, i = 60 deg, r_o = 1000, 64 x 64 on fov 12, rk45
    rtol 1e-6, atol 1e-8, float32, horizon_epsilon 1e-3 (the table there:
    558 captured, 3538 escaped).  Rows of the Numba kernel run interpreted
    (glibc math in the working dtype) and of the NumPy integrator.
    """
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)
    from kerrray.geometry import Spacetime
    from kerrray.raytracing.camera import Camera, initial_states

    cam = Camera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=64)
    Y0 = initial_states(cam, Spacetime(1.0, 0.9))
    out: dict[str, np.ndarray] = {"Y0": Y0}
    t0 = time.time()
    chunks = [Y0[i:i + 32] for i in range(0, len(Y0), 32)]
    with ProcessPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(_float32_chunk, [(c, True) for c in chunks]))
    t1 = time.time()
    numba = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    numpy_rows = _float32_chunk((Y0, False))
    for tag, rows in (("numba", numba), ("numpy", numpy_rows)):
        for k, v in rows.items():
            out[f"{tag}_{k}"] = v
    print(f"float32 zamo: interpreted Numba {t1 - t0:.1f}s, NumPy {time.time() - t1:.1f}s, "
          f"captured {int((numba['state'] == 2).sum())}", flush=True)
    np.savez_compressed(os.path.join(HERE, "float32_zamo64.npz"), **out)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--only", choices=["schwarzschild64", "kerr512", "bisection", "zamo", "disk", "edges", "impacts",
                                       "packing", "float32_zamo", "zamo_rows"], default=None)
    args = ap.parse_args()
    for name, fn in (("schwarzschild64", schwarzschild64), ("kerr512", kerr512), ("bisection", bisection), ("zamo", zamo),
                     ("disk", disk), ("edges", edges), ("impacts", impacts), ("packing", packing),
                     ("float32_zamo", float32_zamo), ("zamo_rows", zamo_rows)):
        if args.only in (None, name):
            fn(args.workers)


if __name__ == "__main__":
    main()
