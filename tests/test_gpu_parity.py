"""Parity of the hip backend with the reference, on an MI355X.

The kernel logic itself is checked bit for bit on the CPU
(test_kernel_host_build.py); on the GPU the device math library rounds sin,
cos and pow differently from glibc in a few percent of calls, so the bar here
is the reference's own backend-agreement pattern (tests/test_numba_backend.py:92-177):

* termination states and the captured mask identical;
* accepted/rejected step counts identical for (almost) every photon;
* where the step sequence is identical, final states within 10 x rtol
  (relative, floor 1) -- except, for CAPTURED photons, p_r, t and phi, which
  are ill-conditioned at the capture radius: p_r ~ sqrt(R)/Delta with
  Delta ~ 1e-6, and t, phi diverge logarithmically at the horizon. There
  the smooth product Delta(r) p_r (= +-sqrt(R(r))) is held to 10 x rtol and
  t, phi to 1e-3 relative; the reference's own NumPy and Numba backends
  differ by ~760 x rtol in p_r on the same inputs (tests/golden, the 64 x 64 Schwarzschild image).

The reference numbers come from tests/golden (the Numba kernel source and
the NumPy integrator run on the same Bardeen initial states) and from the CPU
oracle.
"""

from __future__ import annotations

import numpy as np
import pytest
from conftest import options

pytestmark = pytest.mark.gpu


def rel(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    return np.max(np.abs(a64 - b64) / np.maximum(1.0, np.abs(b64)), axis=-1)


def row_differences(Y, ref_Y, ref_state, spin: float = 0.9, mass: float = 1.0):
    """Per row: the largest relative difference of the components that stay finite (for captured photons
    t and phi, which diverge logarithmically at the horizon, are left out and p_r is replaced by Delta p_r),
    and the relative difference of t and phi."""
    a64, b64 = Y.astype(np.float64), ref_Y.astype(np.float64)
    comp = np.abs(a64 - b64) / np.maximum(1.0, np.abs(b64))
    captured = (ref_state == 2) if ref_state is not None else np.zeros(len(comp), bool)
    a = spin * mass

    def delta_pr(y):
        return (y[:, 1] ** 2 - 2.0 * mass * y[:, 1] + a * a) * y[:, 5]

    dp_rel = np.abs(delta_pr(a64) - delta_pr(b64)) / np.maximum(1.0, np.abs(delta_pr(b64)))
    well = comp.copy()
    well[captured, 0] = 0.0
    well[captured, 3] = 0.0
    well[captured, 5] = dp_rel[captured]
    return well.max(axis=1), comp[:, [0, 3]].max(axis=1), comp, dp_rel, captured


def report(tag: str, mine, ref_Y, ref_steps, ref_rej, rtol: float, ref_state=None, spin: float = 0.9,
           mass: float = 1.0) -> tuple[np.ndarray, float]:
    same_seq = (mine.n_steps == ref_steps) & (mine.n_rejected == ref_rej)
    bitwise = np.all(mine.Y == ref_Y, axis=1)
    r_well, r_log, comp, dp_rel, captured = row_differences(mine.Y, ref_Y, ref_state, spin, mass)
    sel = same_seq & captured
    print(f"[{tag}] photons {len(comp)}  same step sequence {same_seq.mean():.5f}  bitwise {bitwise.mean():.5f}  "
          f"max rel (same seq) {r_well[same_seq].max() if same_seq.any() else 0:.3e}  "
          f"captured p_r {comp[sel, 5].max() if sel.any() else 0:.3e} Delta*p_r {dp_rel[sel].max() if sel.any() else 0:.3e} "
          f"t/phi {r_log[sel].max() if sel.any() else 0:.3e}", flush=True)
    assert np.all(r_well[same_seq] <= 10.0 * rtol), (tag, r_well[same_seq].max())
    assert np.all(r_log[sel] <= 1e-3), (tag, r_log[sel].max())
    return same_seq, float(bitwise.mean())


def test_schwarzschild_64_image(hip, golden1) -> None:
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    integ, term = options(rtol=1e-8, atol=1e-10)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=64)
    img = compute_shadow(Spacetime(1.0, 0.0), cam, integ, term, backend=hip)
    b2 = img.alpha**2 + img.beta**2
    assert np.array_equal(img.captured, b2 < 27.0), "closed-form Schwarzschild mask"
    assert np.array_equal(img.result.state, golden1["numba_state"])
    assert img.n_failed == 0 and img.n_other == 0
    same, _ = report("schwarzschild 64^2", img.result, golden1["numba_Y"], golden1["numba_n_steps"],
                     golden1["numba_n_rejected"], 1e-8, golden1["numba_state"], spin=0.0)
    assert same.mean() >= 0.99
    assert np.all(img.result.max_energy_drift == 0.0) and np.all(img.result.max_lz_drift == 0.0)


def test_kerr_512_mask_is_bit_identical(hip, golden2) -> None:
    from kerrray_amd.camera import BardeenCamera
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    integ, term = options(rtol=1e-6, atol=1e-8)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
    img = compute_shadow(Spacetime(1.0, 0.9), cam, integ, term, backend=hip)
    ref_state = golden2["numpy_state"].astype(np.int64)
    diff = np.flatnonzero(img.result.state != ref_state)
    print(f"[kerr 512^2] kernel {hip.last_kernel_ms:.3f} ms, {len(diff)} pixels differ from the reference mask", flush=True)
    assert np.array_equal(img.result.state, ref_state)
    assert np.array_equal(img.captured[::-1, :], img.captured)
    steps_same = np.mean(img.result.n_steps == golden2["numpy_n_steps"])
    print(f"[kerr 512^2] n_steps identical to the reference on {steps_same:.5f} of pixels", flush=True)
    assert steps_same >= 0.99
    pick = golden2["pick"]
    sub = type("R", (), {})()
    sub.Y, sub.n_steps, sub.n_rejected = img.result.Y[pick], img.result.n_steps[pick], img.result.n_rejected[pick]
    same, _ = report("kerr 512^2 band+sample", sub, golden2["numba_Y"], golden2["numba_n_steps"],
                     golden2["numba_n_rejected"], 1e-6)  # every class within 10 x rtol
    assert same.mean() >= 0.99
    lam_rel = np.abs(img.result.lam[pick] - golden2["numba_lam"]) / np.maximum(1.0, golden2["numba_lam"])
    assert np.all(lam_rel[same] <= 1e-5)


def test_y0_path_equals_image_path(hip) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-7, atol=1e-9)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=8.0, resolution=24)
    img = compute_shadow(st, cam, integ, term, backend=hip)
    batch = hip.integrate_batch(st, bardeen_initial_states(cam, st), integ, term)
    for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift"):
        assert np.array_equal(getattr(img.result, name), getattr(batch, name)), name


def test_paired_rows_equal_single_lane_rows(hip) -> None:
    """Launches of up to 16384 rows carry each photon on a lane pair (the even lane's sincos(theta), the odd
    lane's sincos(2 theta), exchanged by DPP); one row more takes the one-lane kernel. The rows they share
    are equal bit for bit, and so are the point entry's."""
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-7, atol=1e-9)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=128), st)
    paired = hip.integrate_batch(st, Y0, integ, term)
    single = hip.integrate_batch(st, np.concatenate([Y0, Y0[:1]]), integ, term)
    assert len(Y0) == 16384
    for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift"):
        assert np.array_equal(getattr(paired, name), getattr(single, name)[:16384]), name
    print(f"[pairs] 16384 rows: {int((paired.state == 2).sum())} captured, bitwise equal to one lane each",
          flush=True)


def test_batch_matches_oracle_kerr(hip, oracle) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-8, atol=1e-10)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=8.0, resolution=32), st)
    mine = hip.integrate_batch(st, Y0, integ, term)
    ref = oracle.integrate_batch(st, Y0, integ, term)
    assert np.array_equal(mine.state, ref.state)
    same, bitwise = report("oracle a=0.9 rtol 1e-8", mine, ref.Y, ref.n_steps, ref.n_rejected, 1e-8, ref.state)
    assert same.mean() >= 0.99
    esc = same & (ref.state == 1)
    for name in ("max_null_error", "max_carter_drift"):
        a, b = getattr(mine, name)[esc], getattr(ref, name)[esc]
        assert np.all(np.abs(a - b) <= 1e-3 * np.maximum(np.abs(b), 1e-15)), name


def test_float32_variant_matches_oracle(hip, oracle) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-6, atol=1e-8, dtype="float32", horizon_epsilon=1e-3)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=16), st)
    mine = hip.integrate_batch(st, Y0, integ, term)
    ref = oracle.integrate_batch(st, Y0, integ, term)
    # float32 at rtol 1e-6 is round-off dominated (the error norm is mostly float32 noise), so a
    # 1-ulp difference in powf/sinf moves accept/reject decisions; the states, the codes and the
    # (cast-only) energy/L_z drifts are what must agree. The logic is bitwise-checked on the CPU.
    assert mine.Y.dtype == np.float32 and mine.lam.dtype == np.float64
    assert np.array_equal(mine.state, ref.state)
    same = (mine.n_steps == ref.n_steps) & (mine.n_rejected == ref.n_rejected)
    print(f"[float32] same step sequence {same.mean():.3f}, max rel over all photons {np.max(rel(mine.Y, ref.Y)):.2e}")
    assert same.mean() >= 0.8
    assert np.array_equal(mine.max_energy_drift, ref.max_energy_drift)
    assert np.array_equal(mine.max_lz_drift, ref.max_lz_drift)


def test_rk4_and_budgets_match_oracle(hip, oracle) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=12.0, resolution=4), st)
    integ4, far = options(method="rk4", step_size=0.05, lambda_max=20.0, escape_radius=1e6)
    a, b = hip.integrate_batch(st, Y0, integ4, far), oracle.integrate_batch(st, Y0, integ4, far)
    assert np.all(a.state == 3) and np.all(a.n_steps == 400)
    assert np.max(rel(a.Y, b.Y)) < 1e-12 and np.array_equal(a.lam, b.lam)
    integ, term = options(rtol=1e-8, atol=1e-10, max_steps=30)
    a, b = hip.integrate_batch(st, Y0, integ, term), oracle.integrate_batch(st, Y0, integ, term)
    assert np.all(a.state == 3) and np.all(a.n_steps == 30)
    same, _ = report("max_steps=30", a, b.Y, b.n_steps, b.n_rejected, 1e-8, b.state)
    assert same.mean() >= 0.75
    integ, term = options(rtol=1e-8, atol=1e-10, lambda_max=500.0)
    a, b = hip.integrate_batch(st, Y0, integ, term), oracle.integrate_batch(st, Y0, integ, term)
    assert np.all(a.lam == 500.0) and np.array_equal(a.state, b.state)
    same, _ = report("lambda_max=500", a, b.Y, b.n_steps, b.n_rejected, 1e-8, b.state)
    assert same.mean() >= 0.75
    integ, term = options(rtol=1e-8, atol=1e-10, h_max=0.5)
    a, b = hip.integrate_batch(st, Y0, integ, term), oracle.integrate_batch(st, Y0, integ, term)
    assert np.array_equal(a.state, b.state)
    same, _ = report("h_max=0.5", a, b.Y, b.n_steps, b.n_rejected, 1e-8, b.state)
    assert same.mean() >= 0.75


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_disk_event_matches_oracle(hip, oracle, dtype: str) -> None:
    """States crossing theta = pi/2 inside [r_in, r_out] stop as DISK_HIT (numba_kernel.py:439-454)."""
    from kerrray_amd.geodesics import EventOptions
    from kerrray_amd.geometry import Spacetime

    import math

    st = Spacetime(1.0, 0.9)
    # The rays of the reference's test (r = 200, theta = 1.2, E = 1, Q = 15, L_z in {-6, 3, 6, 9},
    # inward, p_theta > 0), written as image points of an observer at theta_o = 1.2.
    th = 1.2
    lz = np.array([-6.0, 3.0, 6.0, 9.0])
    alpha = -lz / math.sin(th)
    beta = -np.sqrt(15.0 + 0.81 * math.cos(th) ** 2 - lz**2 / math.tan(th) ** 2)
    Y0 = oracle.bardeen_points(1.0, 0.9, 200.0, math.degrees(th), alpha, beta)
    events = EventOptions(disk_plane=True, r_in=6.0, r_out=30.0)
    integ, far = options(rtol=1e-7, atol=1e-9, lambda_max=400.0, dtype=dtype, escape_radius=1e6)
    a = hip.integrate_batch(st, Y0, integ, far, events=events)
    b = oracle.integrate_batch(st, Y0, integ, far, events=events)
    assert np.array_equal(a.state, b.state) and np.array_equal(a.event_hit, b.event_hit)
    assert a.event_hit.any()
    hit = a.event_hit
    tol = 1e-9 if dtype == "float64" else 5e-3  # float32: linear crossing interpolant of a different step
    assert np.max(rel(a.event_Y[hit], b.event_Y[hit])) < tol
    assert np.max(np.abs(a.lam - b.lam) / np.maximum(1.0, b.lam)) < tol


def test_schwarzschild_limit_matches_oracle(hip, oracle) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.0)
    integ, term = options(rtol=1e-9, atol=1e-11)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=89.0, fov=8.0, resolution=6), st)
    a, b = hip.integrate_batch(st, Y0, integ, term), oracle.integrate_batch(st, Y0, integ, term)
    assert np.array_equal(a.state, b.state)
    same, _ = report("schwarzschild i=89 rtol 1e-9", a, b.Y, b.n_steps, b.n_rejected, 1e-9, b.state, spin=0.0)
    assert same.mean() >= 0.9


def test_input_validation(hip) -> None:
    from kerrray_amd.geometry import Spacetime

    st = Spacetime(1.0, 0.9)
    integ, term = options()
    with pytest.raises(ValueError):
        hip.integrate_batch(st, np.ones((2, 7)), integ, term)
    with pytest.raises(ValueError):
        hip.integrate_batch(st, np.full((1, 8), np.nan), integ, term)
    with pytest.raises(ValueError):
        hip.integrate_batch(st, np.ones((1, 8)), options(method="dop853")[0], term)
    assert hip.warm_up(st, integ, term) > 0.0
    empty = hip.integrate_batch(st, np.zeros((0, 8)), integ, term)
    assert empty.Y.shape == (0, 8) and empty.state.shape == (0,)


def test_chunked_equals_unchunked(hip) -> None:
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.rays import trace_rays

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-7, atol=1e-9)
    Y0 = bardeen_initial_states(BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=8.0, resolution=8), st)
    fractions: list[float] = []
    one = trace_rays(st, Y0, integ, term, backend=hip)
    chunked = trace_rays(st, Y0, integ, term, backend=hip, chunk_size=20, progress=fractions.append)
    assert fractions == [20 / 64, 40 / 64, 60 / 64, 1.0]
    for name in ("Y", "state", "n_steps", "n_rejected", "lam", "max_null_error", "max_carter_drift"):
        assert np.array_equal(getattr(one, name), getattr(chunked, name)), name


def test_edge_bisection_reproduces_the_reference_brackets(hip, golden_bisect) -> None:
    """study_edge_bisection.py:53-68 semantics: 16 directions x 34 halvings, a = 0.9, i = 60, rtol 1e-6."""
    from kerrray_amd.bisection import bisect_edge
    from kerrray_amd.geometry import Spacetime

    g = golden_bisect
    r_ana = g["r_analytic"]
    res = bisect_edge(Spacetime(1.0, 0.9), 60.0, 1000.0, 1e-6, tuple(g["centre"]), g["angles"], 0.98 * r_ana,
                      1.02 * r_ana, int(g["iterations"]), hip)
    mid_ref = 0.5 * (g["lo"] + g["hi"])
    print(f"[bisect] |gpu - reference| max {np.max(np.abs(res.radius - mid_ref)):.3e} M, "
          f"|gpu - Bardeen| max {np.max(np.abs(res.radius - r_ana)):.3e} M, {res.kernel_ms:.1f} ms", flush=True)
    assert res.failures == 0 and res.nonmonotone == 0
    assert np.max(res.width) <= 1.2e-11
    # Probes in the last rounds sit within the integrator's noise of the edge (~ rtol): a
    # libm-level difference may classify one differently, moving that edge by < 2e-5 M.
    assert np.max(np.abs(res.radius - mid_ref)) < 2e-5
    assert np.max(np.abs(res.radius - r_ana)) < 3e-5


@pytest.mark.parametrize("n_probe", [1, 63])
def test_shadow_edge_matches_bardeen_curve(hip, n_probe: int) -> None:
    """720 directions, rtol 1e-8: edge error of order 2 x rtol (the reference's README:31-32)."""
    from kerrray_amd.bisection import shadow_edge
    from kerrray_amd.geometry import Spacetime

    res, angles, centre, err = shadow_edge(Spacetime(1.0, 0.9), 60.0, 1000.0, 1e-8, 720, backend=hip, n_probe=n_probe)
    print(f"[edge n_probe={n_probe}] {err} width {res.width.max():.2e} {res.kernel_ms:.1f} ms", flush=True)
    assert res.failures == 0 and res.nonmonotone == 0 and np.max(res.width) <= 1e-11
    assert err["rms"] < 5e-8 and err["max"] < 2e-7


@pytest.mark.parametrize("n_probe,n_angles,rounds", [(127, 400, 2), (1, 96, 12), (63, 200, 3), (15, 300, 4),
                                                     (200, 40, 2)])
def test_fused_bisection_rounds_equal_one_call_per_round(hip, n_probe, n_angles, rounds) -> None:
    """Every round in one launch (brackets in LDS between rounds) against one call per round: same brackets bit for bit.

    The cases cover workgroups of 64, 128 and 256 lanes and one to 64
    directions per wavefront.
    """
    from kerrray_amd.bisection import bisect_edge, bisection_options
    from kerrray_amd.boundary import polygon_centroid, polygon_radii
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.orbits import shadow_curve

    st, incl, rtol = Spacetime(1.0, 0.9), 60.0, 1e-8
    integ, term = bisection_options(rtol, 1000.0)
    curve = shadow_curve(st, incl, 4000)
    centre = polygon_centroid(*curve)
    angles = 2.0 * np.pi * np.arange(n_angles) / n_angles
    r_ana = polygon_radii(curve[0], curve[1], centre, angles)
    kw = dict(n_probe=n_probe, integ=integ, term=term)
    fused = bisect_edge(st, incl, 1000.0, rtol, centre, angles, 0.98 * r_ana, 1.02 * r_ana, rounds, hip, **kw)
    lo, hi, failures, nonmono, ms = 0.98 * r_ana, 1.02 * r_ana, 0, 0, 0.0
    for _ in range(rounds):
        one = bisect_edge(st, incl, 1000.0, rtol, centre, angles, lo, hi, 1, hip, **kw)
        lo, hi = one.lo, one.hi
        failures, nonmono, ms = failures + one.failures, nonmono + one.nonmonotone, ms + one.kernel_ms
    print(f"[bisect rounds n_probe={n_probe} x {n_angles}] {rounds} rounds in one launch "
          f"{fused.kernel_ms:.2f} ms, one call per round {ms:.2f} ms", flush=True)
    assert np.array_equal(fused.lo, lo) and np.array_equal(fused.hi, hi)
    assert (fused.failures, fused.nonmonotone) == (failures, nonmono)
    assert np.all(fused.hi - fused.lo < 0.04 * r_ana / (n_probe + 1) ** (rounds - 1))


def test_bisection_edge_cases(hip) -> None:
    """No rounds leaves the brackets as given; the largest one-launch workgroup (256 probes) and the first
    per-round width (257 probes) both narrow every bracket around Bardeen's curve."""
    from kerrray_amd.bisection import bisect_edge
    from kerrray_amd.boundary import polygon_centroid, polygon_radii
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.orbits import shadow_curve

    st, incl = Spacetime(1.0, 0.9), 60.0
    curve = shadow_curve(st, incl, 4000)
    centre = polygon_centroid(*curve)
    angles = 2.0 * np.pi * np.arange(24) / 24
    r_ana = polygon_radii(curve[0], curve[1], centre, angles)
    lo0, hi0 = 0.98 * r_ana, 1.02 * r_ana
    none = bisect_edge(st, incl, 1000.0, 1e-8, centre, angles, lo0, hi0, 0, hip, n_probe=63)
    assert np.array_equal(none.lo, lo0) and np.array_equal(none.hi, hi0) and none.failures == 0
    for n_probe in (256, 257):
        res = bisect_edge(st, incl, 1000.0, 1e-8, centre, angles, lo0, hi0, 2, hip, n_probe=n_probe)
        print(f"[bisect n_probe={n_probe}] width {res.width.max():.2e}, |r - Bardeen| max "
              f"{np.max(np.abs(res.radius - r_ana)):.2e}, {res.kernel_ms:.2f} ms", flush=True)
        assert res.failures == 0 and res.nonmonotone == 0
        assert np.all(res.width <= 0.04 * r_ana / (n_probe + 1) ** 2 * 1.0001)
        assert np.all((res.lo <= r_ana * (1 + 1e-6)) & (res.hi >= r_ana * (1 - 1e-6)))


def test_kerr_512_every_photon_against_the_oracle(hip, oracle, golden2) -> None:
    """All 262,144 photons of the 512^2, a=0.9, i=60 deg, rtol 1e-6 image, device against the oracle.

    * Every class equals the oracle's (and the reference's golden mask).
    * Photons with the oracle's step sequence: final rows within 10 x rtol.
    * Photons whose step sequence differs (a controller decision flipped by a
      last-bit difference of sin/cos): their escape directions at infinity
      (the final state carried to r = infinity by the same device tail for
      both rows) are compared with those of an rtol 1e-12 trace. The
      device's error against that trace must be no larger than four times
      the oracle's (largest and median over those photons): both are valid
      rtol 1e-6 integrations, off the exact geodesic by the same order.
    """
    import os

    import torch

    from kerrray_amd.adjoint import directions_at_infinity
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-6, atol=1e-8)
    cam = BardeenCamera(radius=1000.0, inclination_deg=60.0, fov=10.0, resolution=512)
    img = compute_shadow(st, cam, integ, term, backend=hip)
    Y0 = bardeen_initial_states(cam, st)
    ref = oracle.integrate_batch(st, Y0, integ, term, n_threads=min(16, os.cpu_count() or 1), dynamic=True)
    mine = img.result
    assert np.array_equal(mine.state, ref.state)
    assert np.array_equal(mine.state, golden2["numpy_state"].astype(np.int64))
    same, _ = report("kerr 512^2 every photon", mine, ref.Y, ref.n_steps, ref.n_rejected, 1e-6, ref.state)
    diff = np.flatnonzero(~same)
    esc = diff[ref.state[diff] == 1]
    print(f"  {diff.size} photons with another step sequence, {esc.size} of them escaped", flush=True)
    if esc.size == 0:
        return
    tight_integ, _ = options(rtol=1e-12, atol=1e-14)
    tight = hip.integrate_batch(st, Y0[esc], tight_integ, term)
    assert np.all(tight.state == 1)

    def directions(Y, state):
        th, ph = directions_at_infinity(hip, st, integ, term, torch.from_numpy(np.ascontiguousarray(Y)).cuda(),
                                        torch.from_numpy(np.ascontiguousarray(state)).cuda())
        return np.stack([th.cpu().numpy(), ph.cpu().numpy()], axis=1)

    d_mine = directions(mine.Y[esc], mine.state[esc])
    d_ref = directions(ref.Y[esc], ref.state[esc])
    d_true = directions(tight.Y, tight.state)
    e_mine = np.abs(d_mine - d_true).max(axis=1)
    e_ref = np.abs(d_ref - d_true).max(axis=1)
    print(f"  escape directions against rtol 1e-12: device max {e_mine.max():.2e} median {np.median(e_mine):.2e}; "
          f"oracle max {e_ref.max():.2e} median {np.median(e_ref):.2e}; device vs oracle max "
          f"{np.abs(d_mine - d_ref).max():.2e}", flush=True)
    assert e_mine.max() <= 4.0 * e_ref.max() + 1e-12
    assert np.median(e_mine) <= 4.0 * np.median(e_ref) + 1e-12


def test_masks_equal_the_oracle_on_random_images(hip, oracle) -> None:
    """16 seeded random Bardeen images (spin up to 0.999, inclinations 5-89 deg, 48^2 to 64^2, fields of view
    6-14 M, rtol 1e-5 to 1e-9): compute_shadow's FP64 path (longest-first schedule) gives every pixel the
    termination code of the oracle, the reference kernel restated and run on the host."""
    from kerrray_amd.camera import BardeenCamera, bardeen_initial_states
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.shadow import compute_shadow

    rng = np.random.Generator(np.random.PCG64(13))
    for k in range(16):
        spin = float(rng.uniform(0.0, 0.999))
        incl = float(rng.uniform(5.0, 89.0))
        res = int(rng.choice([48, 56, 64]))
        fov = float(rng.choice([6.0, 10.0, 14.0]))
        rtol = float(10.0 ** -int(rng.integers(5, 10)))
        st = Spacetime(1.0, spin)
        integ, term = options(rtol=rtol, atol=1e-2 * rtol)
        cam = BardeenCamera(radius=1000.0, inclination_deg=incl, fov=fov, resolution=res)
        print(f"[random image {k}] a={spin:.3f} i={incl:.1f} {res}^2 fov {fov} rtol {rtol:g}", flush=True)
        img = compute_shadow(st, cam, integ, term, backend=hip)
        ref = oracle.integrate_batch(st, bardeen_initial_states(cam, st), integ, term, n_threads=8)
        assert np.array_equal(img.result.state, ref.state), k
