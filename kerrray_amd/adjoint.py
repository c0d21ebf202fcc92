"""Differentiation form: gradients of the escape direction by the discrete adjoint of DP5.

The reference has no derivatives of its ray map. Here the map is:
- the Bardeen initial state at the observer, y_0(alpha, beta, a, theta_o)
  (benchmarks/rayset.py:158-165, geodesics/initial_conditions.py:50-112);
- then the accepted Dormand-Prince 5(4) steps of the FP64 adaptive
  integration (raytracing/backends/numba_kernel.py:201-238, 340-463), with
  every accepted step size frozen;
- then the geodesic carried from the final state to r = infinity.

Its outputs are the escape direction (theta_inf, phi_inf) of every escaping
photon. csrc/kr_adjoint.hip records the steps on the device
(``kr_tape_create``) and runs the vector-Jacobian products (``kr_tape_vjp``):
- ``segment=0`` (default): every accepted step's stage states are recorded and
  turned into the step's Jacobian, one step per lane; every photon's 2 x 4
  Jacobian is accumulated through them once, and a product is J^T g
  (632 B of device memory per step while building, 64 B per photon kept);
- ``segment>=1``: step sizes and a checkpoint every ``segment`` steps
  (56 B per step); each product recomputes and sweeps the steps backwards.

* :class:`EscapeDirection` is a ``torch.autograd.Function``: (alpha, beta,
  spin, theta_o) -> (theta_inf, phi_inf, escaped), whose backward is one
  adjoint sweep.
* :func:`lensing_jacobian` returns d(theta_inf, phi_inf)/d(alpha, beta,
  spin, theta_o) per photon (two sweeps).
* :class:`EdgeRadius` and :func:`fit_spin_inclination` fit (spin, theta_o)
  to the shadow edge along fixed directions, by gradient descent. The edge
  radius rho_k along direction psi_k is located by bisection. Its
  derivative comes from the escape-direction Jacobian at the escaping end
  of the final bracket. There, the direction d(rho, p) diverges like
  log(rho - rho_k(p)), so d rho_k / dp = -(d_rho d . d_p d) / |d_rho d|^2,
  up to terms of the order of the bracket width.

theta_o is in radians throughout. Gradients are float64 on the device of
the inputs.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from kerrray_amd import _lib
from kerrray_amd.backends import HipBackend, get_backend, kernel_params
from kerrray_amd.bisection import bisect_edge, bisection_options, rounds_for_width
from kerrray_amd.geodesics import IntegratorOptions, TerminationOptions
from kerrray_amd.geometry import Spacetime
from kerrray_amd.photons import TerminationState

__all__ = [
    "EdgeFit", "EdgeRadius", "EscapeDirection", "Tape", "directions_at_infinity", "edge_radii",
    "escape_directions", "fit_spin_inclination", "lensing_jacobian",
]

DEFAULT_SEGMENT = 0  # Jacobian tape; >= 1: checkpoints every `segment` steps
TAPE_FLAG_REPLAY = 1 << 8  # a re-integration differed from the trace
TAPE_FLAG_RECOMPUTE = 1 << 9  # a recomputed final state differed from the traced one


def _engine(backend) -> HipBackend:
    engine = get_backend(backend) if isinstance(backend, str) else backend
    if not isinstance(engine, HipBackend):
        raise TypeError("the adjoint runs on the hip backend")
    return engine


def _outputs(n: int, device: torch.device) -> tuple[dict[str, torch.Tensor], _lib.KrOutputs]:
    f64 = dict(dtype=torch.float64, device=device)
    i64 = dict(dtype=torch.int64, device=device)
    bufs = {"Y": torch.empty((n, 8), **f64), "lam": torch.empty(n, **f64), "n_steps": torch.empty(n, **i64),
            "n_rejected": torch.empty(n, **i64), "state": torch.empty(n, **i64),
            "m0": torch.empty(n, **f64), "m1": torch.empty(n, **f64), "m2": torch.empty(n, **f64),
            "m3": torch.empty(n, **f64)}
    o = _lib.KrOutputs()
    o.Y, o.lam, o.n_steps, o.n_rejected, o.state = (bufs[k].data_ptr() for k in
                                                    ("Y", "lam", "n_steps", "n_rejected", "state"))
    o.max_null_error, o.max_energy_drift, o.max_lz_drift, o.max_carter_drift = (
        bufs[k].data_ptr() for k in ("m0", "m1", "m2", "m3"))
    return bufs, o


def directions_at_infinity(backend, st: Spacetime, integ: IntegratorOptions, term: TerminationOptions,
                           Y: torch.Tensor, state: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(theta_inf, phi_inf) of FP64 result rows Y (n, 8) on the device; NaN where state is not ESCAPED."""
    engine = _engine(backend)
    Y = Y.to(torch.float64).contiguous()
    state = state.to(torch.int64).contiguous()
    n = int(state.numel())
    th = torch.empty(n, dtype=torch.float64, device=Y.device)
    ph = torch.empty_like(th)
    params = kernel_params(st, integ, term)
    stream = torch.cuda.current_stream(Y.device).cuda_stream
    _lib.check(_lib.lib().kr_escape_directions_dev(engine.handle.ptr, ctypes.byref(params), Y.data_ptr(),
                                                   state.data_ptr(), n, th.data_ptr(), ph.data_ptr(),
                                                   ctypes.c_void_p(stream)), "kr_escape_directions_dev")
    return th, ph


class Tape:
    """The recorded FP64 trace of image-plane points, ready for adjoint sweeps.

    ``alpha``/``beta`` are float64 CUDA tensors; ``spin`` and ``theta_o``
    (radians) are numbers.
    """

    def __init__(self, engine: HipBackend, alpha: torch.Tensor, beta: torch.Tensor, spin: float, theta_o: float, *,
                 integ: IntegratorOptions, term: TerminationOptions, mass: float = 1.0, r_o: float = 1000.0,
                 segment: int = DEFAULT_SEGMENT) -> None:
        if alpha.shape != beta.shape or alpha.dim() != 1:
            raise ValueError("alpha and beta must be 1-D tensors of one shape")
        if integ.method != "rk45" or np.dtype(integ.dtype) != np.float64:
            raise ValueError("the adjoint is implemented for FP64 Dormand-Prince 5(4)")
        self.engine = engine
        self.alpha = alpha.detach().to(torch.float64).contiguous()
        self.beta = beta.detach().to(torch.float64).contiguous()
        if not self.alpha.is_cuda:
            raise ValueError("alpha and beta must be CUDA tensors")
        n = int(self.alpha.numel())
        self.n = n
        st = Spacetime(float(mass), float(spin))
        self.params = kernel_params(st, integ, term)
        self.image = _lib.KrImage(float(mass), float(spin), float(r_o), float(theta_o), 1.0, 1)
        self.bufs, self.outputs = _outputs(n, self.alpha.device)
        self.theta_inf = torch.empty(n, dtype=torch.float64, device=self.alpha.device)
        self.phi_inf = torch.empty(n, dtype=torch.float64, device=self.alpha.device)
        torch.cuda.current_stream(self.alpha.device).synchronize()
        tape = ctypes.c_void_p()
        with engine.handle._lock:
            _lib.check(
                _lib.lib().kr_tape_create(
                    engine.handle.ptr, ctypes.byref(self.params), ctypes.byref(self.image), self.alpha.data_ptr(),
                    self.beta.data_ptr(), n, int(segment), ctypes.byref(self.outputs), self.theta_inf.data_ptr(),
                    self.phi_inf.data_ptr(), ctypes.byref(tape)),
                "kr_tape_create",
            )
        self._tape = tape
        self.state = self.bufs["state"]
        self.escaped = self.state == int(TerminationState.ESCAPED)
        self.recorded = int(_lib.lib().kr_tape_recorded(tape))
        self.steps = int(_lib.lib().kr_tape_steps(tape))
        self.record_ms = float(_lib.lib().kr_tape_record_ms(tape))
        self.jacobian_ms = float(_lib.lib().kr_tape_jacobian_ms(tape))

    def vjp(self, g_theta: torch.Tensor, g_phi: torch.Tensor) -> torch.Tensor:
        """(n, 4): g_theta d theta_inf/dx + g_phi d phi_inf/dx, x = (alpha, beta, spin, theta_o)."""
        dev = self.alpha.device
        gt = torch.where(self.escaped, g_theta.to(dev, torch.float64), 0.0).contiguous()
        gp = torch.where(self.escaped, g_phi.to(dev, torch.float64), 0.0).contiguous()
        grad = torch.empty((self.n, 4), dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(_lib.lib().kr_tape_vjp(self._tape, gt.data_ptr(), gp.data_ptr(), grad.data_ptr(),
                                          ctypes.c_void_p(stream)), "kr_tape_vjp")
        return grad

    def jacobian(self) -> torch.Tensor:
        """(n, 2, 4): d(theta_inf, phi_inf)/d(alpha, beta, spin, theta_o); 0 rows for non-escaped photons."""
        one = torch.ones(self.n, dtype=torch.float64, device=self.alpha.device)
        zero = torch.zeros_like(one)
        return torch.stack([self.vjp(one, zero), self.vjp(zero, one)], dim=1)

    def flags(self) -> int:
        f = ctypes.c_int32(0)
        _lib.check(_lib.lib().kr_tape_flags(self._tape, ctypes.byref(f)), "kr_tape_flags")
        return int(f.value)

    def close(self) -> None:
        if getattr(self, "_tape", None) is not None and self._tape.value:
            _lib.lib().kr_tape_destroy(self._tape)
            self._tape = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


@dataclass(frozen=True)
class TapeOptions:
    integ: IntegratorOptions
    term: TerminationOptions
    mass: float = 1.0
    r_o: float = 1000.0
    segment: int = DEFAULT_SEGMENT


class EscapeDirection(torch.autograd.Function):
    """(alpha, beta, spin, theta_o) -> (theta_inf, phi_inf, escaped).

    ``alpha``, ``beta``: (n,) float64 CUDA tensors; ``spin``, ``theta_o``:
    0-d float64 tensors. theta_inf/phi_inf are NaN, and receive no gradient,
    for photons that did not escape.
    """

    @staticmethod
    def forward(ctx, alpha, beta, spin, theta_o, engine, opts: TapeOptions):
        tape = Tape(engine, alpha, beta, float(spin), float(theta_o), integ=opts.integ, term=opts.term,
                    mass=opts.mass, r_o=opts.r_o, segment=opts.segment)
        ctx.tape = tape
        ctx.spin_like = (spin.shape, spin.dtype, spin.device)
        ctx.theta_like = (theta_o.shape, theta_o.dtype, theta_o.device)
        escaped = tape.escaped.clone()
        ctx.mark_non_differentiable(escaped)
        return tape.theta_inf.clone(), tape.phi_inf.clone(), escaped

    @staticmethod
    def backward(ctx, g_theta, g_phi, _g_escaped):
        tape: Tape = ctx.tape
        zero = torch.zeros(tape.n, dtype=torch.float64, device=tape.alpha.device)
        g_theta = zero if g_theta is None else torch.nan_to_num(g_theta, nan=0.0)
        g_phi = zero if g_phi is None else torch.nan_to_num(g_phi, nan=0.0)
        grad = tape.vjp(g_theta, g_phi)
        g_spin = grad[:, 2].sum().reshape(ctx.spin_like[0]).to(ctx.spin_like[1]).to(ctx.spin_like[2])
        g_th = grad[:, 3].sum().reshape(ctx.theta_like[0]).to(ctx.theta_like[1]).to(ctx.theta_like[2])
        return grad[:, 0], grad[:, 1], g_spin, g_th, None, None


def escape_directions(alpha: torch.Tensor, beta: torch.Tensor, spin: torch.Tensor, theta_o: torch.Tensor, *,
                      integ: IntegratorOptions, term: TerminationOptions, backend="hip", mass: float = 1.0,
                      r_o: float = 1000.0, segment: int = DEFAULT_SEGMENT):
    """Differentiable (theta_inf, phi_inf, escaped) of the image-plane points (alpha, beta)."""
    opts = TapeOptions(integ, term, mass, r_o, segment)
    return EscapeDirection.apply(alpha, beta, spin, theta_o, _engine(backend), opts)


def lensing_jacobian(alpha: torch.Tensor, beta: torch.Tensor, spin: float, theta_o: float, *,
                     integ: IntegratorOptions, term: TerminationOptions, backend="hip", mass: float = 1.0,
                     r_o: float = 1000.0, segment: int = DEFAULT_SEGMENT):
    """(theta_inf, phi_inf, escaped, J) with J (n, 2, 4) = d(theta_inf, phi_inf)/d(alpha, beta, spin, theta_o)."""
    tape = Tape(_engine(backend), alpha, beta, spin, theta_o, integ=integ, term=term, mass=mass, r_o=r_o,
                segment=segment)
    try:
        J = tape.jacobian()
        flags = tape.flags()
        if flags & (TAPE_FLAG_REPLAY | TAPE_FLAG_RECOMPUTE):
            raise RuntimeError(f"adjoint tape consistency flags {flags:#x}")
        return tape.theta_inf, tape.phi_inf, tape.escaped, J
    finally:
        tape.close()


# ----------------------------------------------------------------------------- edge fit


@dataclass(frozen=True)
class EdgeSetup:
    """Directions psi_k about a fixed centre, the bracket and the integrator of an edge fit.

    The centre must lie inside the shadow for every (spin, theta_o) the fit
    visits, and off the line alpha = 0. A direction running along alpha = 0
    has L_z = 0 for all of its photons. Such photons cross the polar axis,
    which the Boyer-Lindquist classification reports as out of domain.
    """

    angles: np.ndarray
    centre: tuple[float, float] = (0.5, 0.0)
    r_lo: float = 2.0
    r_hi: float = 9.0
    width: float = 1e-9
    n_probe: int = 63
    rtol: float = 1e-8
    mass: float = 1.0
    r_o: float = 1000.0
    segment: int = DEFAULT_SEGMENT


def edge_radii(engine: HipBackend, spin: float, theta_o: float, setup: EdgeSetup, *, with_derivatives: bool = False):
    """Edge radius along each direction (midpoint of the final bracket), and optionally d rho/d(spin, theta_o).

    Returns (rho, drho) with drho (n, 2) or None. The derivative uses the
    escape-direction Jacobian at the escaping end of each final bracket.
    """
    st = Spacetime(setup.mass, spin)
    incl = math.degrees(theta_o)
    rounds = rounds_for_width(setup.r_hi - setup.r_lo, setup.width, setup.n_probe)
    res = bisect_edge(st, incl, setup.r_o, setup.rtol, setup.centre, setup.angles, setup.r_lo, setup.r_hi, rounds,
                      engine, n_probe=setup.n_probe)
    if res.failures:
        raise RuntimeError(f"{res.failures} bisection probes neither escaped nor were captured")
    rho = res.radius
    if not with_derivatives:
        return rho, None
    integ, term = bisection_options(setup.rtol, setup.r_o)
    cos_k, sin_k = np.cos(setup.angles), np.sin(setup.angles)
    dev = torch.device("cuda", engine.handle.device)
    a = torch.as_tensor(setup.centre[0] + res.hi * cos_k, dtype=torch.float64, device=dev)
    b = torch.as_tensor(setup.centre[1] + res.hi * sin_k, dtype=torch.float64, device=dev)
    _, _, escaped, J = lensing_jacobian(a, b, spin, math.radians(incl), integ=integ, term=term, backend=engine,
                                        mass=setup.mass, r_o=setup.r_o, segment=setup.segment)
    if not bool(escaped.all()):
        raise RuntimeError("an escaping bracket end did not escape when re-traced")
    J = J.cpu().numpy()
    d_rho = cos_k[:, None] * J[:, :, 0] + sin_k[:, None] * J[:, :, 1]  # (n, 2)
    norm2 = np.sum(d_rho * d_rho, axis=1)
    drho = -np.einsum("ki,kij->kj", d_rho, J[:, :, 2:4]) / norm2[:, None]
    return rho, drho


class EdgeRadius(torch.autograd.Function):
    """(spin, theta_o) -> edge radii rho_k along the setup's directions."""

    @staticmethod
    def forward(ctx, spin, theta_o, engine, setup: EdgeSetup):
        rho, drho = edge_radii(engine, float(spin), float(theta_o), setup, with_derivatives=True)
        ctx.drho = torch.as_tensor(drho, dtype=torch.float64, device=spin.device)
        return torch.as_tensor(rho, dtype=torch.float64, device=spin.device)

    @staticmethod
    def backward(ctx, g):
        gs = (g * ctx.drho[:, 0]).sum()
        gt = (g * ctx.drho[:, 1]).sum()
        return gs, gt, None, None


@dataclass
class EdgeFit:
    spin: float
    theta_o: float
    loss: float
    iterations: int
    converged: bool
    history: list[tuple[float, float, float]] = field(default_factory=list)  # (spin, theta_o, loss)


def fit_spin_inclination(target: np.ndarray, setup: EdgeSetup, spin0: float, theta_o0: float, *, backend="hip",
                         max_iter: int = 60, gtol: float = 1e-12, xtol: float = 1e-9, step0: float = 1e-2,
                         max_move: float = 0.05) -> EdgeFit:
    """Gradient descent on L = mean_k (rho_k(spin, theta_o) - target_k)^2.

    Step lengths follow Barzilai-Borwein from the last two gradients. A move
    is capped at ``max_move`` in (spin, theta_o) and halved until L
    decreases. A trial point whose bisection has failing photons counts as
    L = inf. spin stays inside (-1, 1) and theta_o inside (0, pi). The
    gradient comes from :class:`EdgeRadius`. Stops when the gradient norm
    falls below ``gtol`` or an accepted step moves both parameters by less
    than ``xtol``.
    """
    engine = _engine(backend)
    tgt = torch.as_tensor(np.asarray(target, dtype=np.float64))
    x = np.array([spin0, theta_o0], dtype=np.float64)

    def loss_grad(v: np.ndarray) -> tuple[float, np.ndarray]:
        s = torch.tensor(v[0], dtype=torch.float64, requires_grad=True)
        t = torch.tensor(v[1], dtype=torch.float64, requires_grad=True)
        rho = EdgeRadius.apply(s, t, engine, setup)
        L = torch.mean((rho - tgt) ** 2)
        L.backward()
        return float(L.detach()), np.array([float(s.grad), float(t.grad)])

    def loss_only(v: np.ndarray) -> float:
        try:
            rho, _ = edge_radii(engine, float(v[0]), float(v[1]), setup)
        except RuntimeError:
            return math.inf
        return float(np.mean((rho - tgt.numpy()) ** 2))

    def clip(v: np.ndarray) -> np.ndarray:
        return np.array([min(max(v[0], -0.999), 0.999), min(max(v[1], 1e-3), math.pi - 1e-3)])

    L, g = loss_grad(x)
    history = [(float(x[0]), float(x[1]), L)]
    step = step0 / max(float(np.linalg.norm(g)), 1e-300)
    x_prev = g_prev = None
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        if np.linalg.norm(g) < gtol:
            converged = True
            break
        if x_prev is not None:
            dx, dg = x - x_prev, g - g_prev
            denom = float(dx @ dg)
            if denom > 0.0:
                step = float(dx @ dx) / denom
        while True:
            d = step * g
            norm = float(np.linalg.norm(d))
            if norm > max_move:
                d *= max_move / norm
            x_new = clip(x - d)
            L_new = loss_only(x_new)
            if L_new < L or norm < 1e-15:
                break
            step *= 0.5
        if not L_new < L:  # no decrease at any step length: at the noise floor
            converged = True
            break
        x_prev, g_prev = x, g
        x = x_new
        L, g = loss_grad(x)
        history.append((float(x[0]), float(x[1]), L))
        if np.max(np.abs(x - x_prev)) < xtol:
            converged = True
            break
    return EdgeFit(spin=float(x[0]), theta_o=float(x[1]), loss=L, iterations=it, converged=converged,
                   history=history)
