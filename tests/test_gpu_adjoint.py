"""Differentiation form on the GPU.

The workload is a 256 x 256 image at a in {0.5, 0.9}, i = 60 deg, rtol 1e-8
(atol 1e-10). The map differentiated is the escape direction at infinity of
every escaping photon, with respect to (alpha, beta, spin, theta_o).

* Against the oracle's forward-mode variational integration, on a strided
  sample. Where the GPU took the oracle's step sequence (same step counts),
  both differentiate the same discrete map up to the 1-ulp differences of
  device sin/cos, which orbits near the critical curve amplify: median
  1e-12, worst 1e-6 relative. Where
  device sin/cos (1 ulp from glibc) moved a step decision, the maps differ
  by a global-error-sized amount and the tolerance is the finite-difference
  one below.
* Against central finite differences (h = 1e-6) of the GPU's own adaptive
  integration. The per-element tolerance follows from integration noise /
  2h, about 1e-8 / 2e-6 at worst, where a step decision flips. It is
  asserted as a median < 1e-6 and a 99th percentile < 1e-3 relative.
* The autograd Function's backward is the same contraction.
* Edge fit: edge-radius derivatives against the analytic Bardeen curve,
  then recovery of (a, theta_o) = (0.7, 40 deg) from a synthetic 360-angle
  edge.
"""

from __future__ import annotations

import math
import time

import numpy as np
import pytest
from conftest import options

pytestmark = pytest.mark.gpu

RES, INCL, RTOL = 256, 60.0, 1e-8


def _grid(torch):
    from kerrray_amd.camera import BardeenCamera

    cam = BardeenCamera(radius=1000.0, inclination_deg=INCL, fov=10.0, resolution=RES)
    a, b = cam.pixel_coordinates()
    dev = torch.device("cuda:0")
    return (torch.as_tensor(np.ascontiguousarray(a.ravel()), device=dev),
            torch.as_tensor(np.ascontiguousarray(b.ravel()), device=dev))


def _tape(hip, alpha, beta, spin, theta_o):
    from kerrray_amd.adjoint import Tape

    integ, term = options(rtol=RTOL, atol=1e-2 * RTOL)
    return Tape(hip, alpha, beta, spin, theta_o, integ=integ, term=term)


@pytest.mark.parametrize("spin", [0.5, 0.9])
def test_jacobian_map_equals_forward_variational_integration(hip, oracle, spin) -> None:
    import torch

    alpha, beta = _grid(torch)
    th_o = math.radians(INCL)
    t0 = time.perf_counter()
    tape = _tape(hip, alpha, beta, spin, th_o)
    J = tape.jacobian().cpu().numpy()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert tape.flags() == 0
    esc = tape.escaped.cpu().numpy()
    print(f"[adjoint 256^2 a={spin}] escaped {esc.sum()}, accepted steps on tape {tape.steps}, "
          f"record {tape.record_ms:.1f} ms, wall {wall:.2f} s", flush=True)
    idx = np.arange(0, RES * RES, 23)
    a_np, b_np = alpha.cpu().numpy()[idx], beta.cpu().numpy()[idx]
    V = oracle.escape_variational(type("S", (), {"mass": 1.0, "spin": spin})(), th_o, 1000.0, a_np, b_np,
                                  *options(rtol=RTOL, atol=1e-2 * RTOL))
    gpu_state = tape.state.cpu().numpy()[idx]
    assert np.mean(gpu_state == V.state) > 0.999
    both = esc[idx] & (V.state == 1)
    same_path = both & (tape.bufs["n_steps"].cpu().numpy()[idx] == _oracle_steps(oracle, spin, a_np, b_np))
    Jg, Jo = J[idx], V.jac
    scale = np.maximum(np.abs(Jo), 1e-9 * np.nanmax(np.abs(Jo[both])))
    rel = np.abs(Jg - Jo) / scale
    th_gpu = tape.theta_inf.cpu().numpy()[idx]
    dth = np.abs(th_gpu - V.theta)
    print(f"  sample {idx.size}: both escaped {both.sum()}, same step count {same_path.sum()}; "
          f"rel err same-path max {rel[same_path].max():.2e}, all median {np.median(rel[both]):.2e} "
          f"p99 {np.quantile(rel[both], 0.99):.2e}; theta_inf |diff| max {dth[both].max():.2e}", flush=True)
    assert same_path.sum() > 0.5 * both.sum()
    # Same step sequence: the same discrete map up to 1-ulp sin/cos
    # differences, which near-critical orbits amplify.
    assert np.median(rel[same_path]) < 1e-12 and rel[same_path].max() < 1e-6
    assert np.median(rel[both]) < 1e-6 and np.quantile(rel[both], 0.99) < 1e-3


@pytest.mark.parametrize("spin", [0.5, 0.9])
def test_jacobian_tape_equals_the_checkpoint_sweep(hip, spin) -> None:
    """The Jacobian tape (segment 0: photon Jacobians accumulated through step Jacobians) and the checkpointed
    reverse sweep (segment 32) differentiate the same discrete map; they differ in the order of their
    roundings only, which near-critical orbits amplify."""
    import torch

    from kerrray_amd.adjoint import Tape

    alpha, beta = _grid(torch)
    integ, term = options(rtol=RTOL, atol=1e-2 * RTOL)
    tapes = [Tape(hip, alpha, beta, spin, math.radians(INCL), integ=integ, term=term, segment=seg) for seg in (0, 32)]
    J0, J1 = (t.jacobian().cpu().numpy() for t in tapes)
    for t in tapes:
        assert t.flags() == 0
    esc = tapes[0].escaped.cpu().numpy()
    assert np.array_equal(esc, tapes[1].escaped.cpu().numpy())
    assert tapes[0].steps == tapes[1].steps
    scale = np.maximum(np.abs(J1[esc]), 1e-9 * np.abs(J1[esc]).max())
    rel = np.abs(J0[esc] - J1[esc]) / scale
    print(f"[Jacobian tape vs checkpoints a={spin}] {esc.sum()} photons, {tapes[0].steps} steps: record "
          f"{tapes[0].record_ms:.1f} ms + Jacobians {tapes[0].jacobian_ms:.1f} ms; rel diff median "
          f"{np.median(rel):.1e} max {rel.max():.1e}", flush=True)
    assert np.median(rel) < 1e-12 and rel.max() < 1e-6
    assert np.all(J0[~esc] == 0.0)
    for t in tapes:
        t.close()


def _oracle_steps(oracle, spin, a, b):
    integ, term = options(rtol=RTOL, atol=1e-2 * RTOL)
    Y0 = oracle.bardeen_points(1.0, spin, 1000.0, INCL, a, b)
    return oracle.integrate_batch(type("S", (), {"mass": 1.0, "spin": spin})(), Y0, integ, term).n_steps


@pytest.mark.parametrize("spin", [0.5, 0.9])
def test_jacobian_map_agrees_with_central_differences(hip, spin) -> None:
    import torch

    alpha, beta = _grid(torch)
    th_o = math.radians(INCL)
    base = _tape(hip, alpha, beta, spin, th_o)
    J = base.jacobian().cpu().numpy()
    ok = base.escaped.cpu().numpy().copy()
    h = 1e-6
    FD = np.empty_like(J)
    for j in range(4):
        args_p = [alpha, beta, spin, th_o]
        args_m = [alpha, beta, spin, th_o]
        args_p[j] = args_p[j] + h
        args_m[j] = args_m[j] - h
        tp, tm = _tape(hip, *args_p), _tape(hip, *args_m)
        ok &= tp.escaped.cpu().numpy() & tm.escaped.cpu().numpy()
        FD[:, 0, j] = ((tp.theta_inf - tm.theta_inf) / (2 * h)).cpu().numpy()
        FD[:, 1, j] = ((tp.phi_inf - tm.phi_inf) / (2 * h)).cpu().numpy()
        tp.close()
        tm.close()
    rel = np.abs(FD[ok] - J[ok]) / np.maximum(np.abs(J[ok]), 1e-3)
    print(f"[adjoint vs FD a={spin}] photons {ok.sum()}: median {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e}, "
          f"max {rel.max():.2e}", flush=True)
    assert ok.sum() > 0.7 * RES * RES
    assert np.median(rel) < 1e-6 and np.quantile(rel, 0.99) < 1e-3


def test_autograd_backward_is_the_jacobian_contraction(hip) -> None:
    import torch

    from kerrray_amd.adjoint import escape_directions

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(5)
    alpha = (torch.rand(512, generator=g, dtype=torch.float64) * 20 - 10).to(dev).requires_grad_(True)
    beta = (torch.rand(512, generator=g, dtype=torch.float64) * 20 - 10).to(dev).requires_grad_(True)
    spin = torch.tensor(0.9, dtype=torch.float64, requires_grad=True)
    th_o = torch.tensor(math.radians(INCL), dtype=torch.float64, requires_grad=True)
    integ, term = options(rtol=RTOL, atol=1e-2 * RTOL)
    th, ph, esc = escape_directions(alpha, beta, spin, th_o, integ=integ, term=term, backend=hip)
    w = torch.rand(512, generator=g, dtype=torch.float64).to(dev)
    v = torch.rand(512, generator=g, dtype=torch.float64).to(dev)
    loss = (torch.where(esc, w * th, 0.0) + torch.where(esc, v * ph, 0.0)).sum()
    loss.backward()
    tape = _tape(hip, alpha.detach(), beta.detach(), 0.9, math.radians(INCL))
    J = tape.jacobian()
    e = esc.to(torch.float64)
    ref = (w * e)[:, None] * J[:, 0, :] + (v * e)[:, None] * J[:, 1, :]
    assert torch.equal(tape.escaped, esc)
    # One sweep seeded with (w, v) against w J_theta + v J_phi from two
    # sweeps: equal up to the rounding of that linear combination.
    torch.testing.assert_close(alpha.grad, ref[:, 0], rtol=1e-10, atol=1e-13 * float(ref[:, 0].abs().max()))
    torch.testing.assert_close(beta.grad, ref[:, 1], rtol=1e-10, atol=1e-13 * float(ref[:, 1].abs().max()))
    torch.testing.assert_close(spin.grad, ref[:, 2].sum().cpu(), rtol=1e-10, atol=0)
    torch.testing.assert_close(th_o.grad, ref[:, 3].sum().cpu(), rtol=1e-10, atol=0)


def test_edge_radius_derivative_matches_the_bardeen_curve(hip) -> None:
    """d rho_k / d(a, theta_o) from the adjoint vs central differences of the analytic curve."""
    from kerrray_amd.adjoint import EdgeSetup, edge_radii
    from kerrray_amd.boundary import polygon_radii
    from kerrray_amd.geometry import Spacetime
    from kerrray_amd.orbits import shadow_curve

    angles = 2 * np.pi * np.arange(72) / 72
    setup = EdgeSetup(angles=angles)
    spin, th_o = 0.7, math.radians(40.0)
    rho, drho = edge_radii(hip, spin, th_o, setup, with_derivatives=True)

    def curve_radii(a, t):
        c = shadow_curve(Spacetime(1.0, a), math.degrees(t), 40000)
        return polygon_radii(c[0], c[1], setup.centre, angles)

    r_ana = curve_radii(spin, th_o)
    h = 1e-5
    d_ana = np.stack([(curve_radii(spin + h, th_o) - curve_radii(spin - h, th_o)) / (2 * h),
                      (curve_radii(spin, th_o + h) - curve_radii(spin, th_o - h)) / (2 * h)], axis=1)
    err = np.abs(drho - d_ana)
    print(f"[edge derivative] |rho - Bardeen| max {np.abs(rho - r_ana).max():.2e}; d rho/da max {np.abs(d_ana[:, 0]).max():.3f} "
          f"err max {err[:, 0].max():.2e}; d rho/dtheta max {np.abs(d_ana[:, 1]).max():.3f} err max {err[:, 1].max():.2e}",
          flush=True)
    assert np.abs(rho - r_ana).max() < 1e-5
    assert np.all(err <= 1e-4 * (1.0 + np.abs(d_ana)))


def test_recover_spin_and_inclination_from_a_synthetic_edge(hip) -> None:
    from kerrray_amd.adjoint import EdgeSetup, edge_radii, fit_spin_inclination

    setup = EdgeSetup(angles=2 * np.pi * np.arange(360) / 360)
    true = (0.7, math.radians(40.0))
    target, _ = edge_radii(hip, *true, setup)
    t0 = time.perf_counter()
    fit = fit_spin_inclination(target, setup, 0.4, math.radians(55.0), backend=hip)
    wall = time.perf_counter() - t0
    print(f"[edge fit] a {fit.spin:.9f} (true 0.7), theta_o {math.degrees(fit.theta_o):.9f} deg (true 40), "
          f"loss {fit.loss:.2e}, iterations {fit.iterations}, converged {fit.converged}, {wall:.1f} s", flush=True)
    assert abs(fit.spin - true[0]) < 1e-5
    assert abs(fit.theta_o - true[1]) < 1e-5
