"""Shadow-edge geometry on the image plane.

Polygon radii and centroid, the sub-pixel edge of a captured mask and the
comparison with an analytic curve, as in src/kerrray/raytracing/boundary.py:74-307:

* ``polygon_radii``: exact ray/segment intersection, smallest t > 0 with
  0 <= s <= 1 (widened by 1e-9 so a ray through a vertex is not lost),
  searching only the segment a ray points into when the polygon is
  star-shaped about the centre;
* ``polygon_centroid``: shoelace area centroid;
* ``extract_boundary_from_mask``: bilinear 0.5 level set of the mask (zero
  outside the image) marched along polar rays in steps of 1/8 pixel, on the
  device (one lane per direction);
* ``edge_error``: radial differences between traced edge radii and the curve.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any

import numpy as np
from numpy.typing import ArrayLike, NDArray

from kerrray_amd.geometry import Spacetime

__all__ = [
    "EDGE_TOLERANCE", "STEP_FRACTION", "edge_error", "extract_boundary_from_mask", "mask_centroid",
    "polygon_centroid", "polygon_radii",
]

Float = NDArray[np.float64]
STEP_FRACTION = 0.125
EDGE_TOLERANCE = 1e-9


def _ray_segment_t(qx: Float, qy: Float, ex: Float, ey: Float, dx: Float, dy: Float) -> Float:
    """Ray parameter t of the intersection of rays (dx, dy) from the origin with segments q + s e (inf where the
    intersection has t <= 0 or s outside [0, 1] widened by EDGE_TOLERANCE); broadcasts."""
    det = dx * (-ey) - dy * (-ex)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (qx * (-ey) - qy * (-ex)) / det
        s = (dx * qy - dy * qx) / det
    valid = np.isfinite(t) & (t > 0.0) & (s >= -EDGE_TOLERANCE) & (s <= 1.0 + EDGE_TOLERANCE)
    return np.where(valid, t, np.inf)


def _star_candidates(qx: Float, qy: Float, phi: Float) -> NDArray[np.int64] | None:
    """For a polygon whose vertices turn monotonically once around the origin (star-shaped about it), the
    segment each ray direction phi points into and its two neighbours, (len(phi), 3); None otherwise."""
    n = qx.size
    va = np.arctan2(qy, qx)
    steps = np.diff(np.unwrap(np.append(va, va[0])))
    turn = steps.sum()
    if np.all(steps > 0.0) and abs(turn - 2.0 * np.pi) < 1e-9:
        sign = 1.0
    elif np.all(steps < 0.0) and abs(turn + 2.0 * np.pi) < 1e-9:
        sign = -1.0
    else:
        return None
    u = sign * np.concatenate([[0.0], np.cumsum(steps)])  # increasing from 0 to 2 pi
    x = np.mod(sign * (phi - va[0]), 2.0 * np.pi)
    j = np.clip(np.searchsorted(u, x, side="right") - 1, 0, n - 1)
    return np.stack([(j - 1) % n, j, (j + 1) % n], axis=1)


def polygon_radii(alpha: ArrayLike, beta: ArrayLike, centre: tuple[float, float], angles: ArrayLike) -> Float:
    """Distance from ``centre`` to the closed polygon along each angle (the smallest valid ray parameter over
    all segments).  For a polygon star-shaped about the centre, only the segment a ray points into and its
    neighbours can hold it, and only those are evaluated (same arithmetic, same result)."""
    px = np.asarray(alpha, dtype=np.float64).ravel()
    py = np.asarray(beta, dtype=np.float64).ravel()
    if px.size < 3 or px.shape != py.shape:
        raise ValueError("a polygon needs at least three (alpha, beta) points")
    phi = np.asarray(angles, dtype=np.float64).ravel()
    qx, qy = px - float(centre[0]), py - float(centre[1])
    ex, ey = np.roll(qx, -1) - qx, np.roll(qy, -1) - qy
    dx, dy = np.cos(phi)[:, None], np.sin(phi)[:, None]
    cand = _star_candidates(qx, qy, phi)
    if cand is not None:
        radii = _ray_segment_t(qx[cand], qy[cand], ex[cand], ey[cand], dx, dy).min(axis=1)
    else:
        radii = _ray_segment_t(qx[None, :], qy[None, :], ex[None, :], ey[None, :], dx, dy).min(axis=1)
    if not np.all(np.isfinite(radii)):
        raise ValueError("a polar ray from the centre does not intersect the curve")
    return radii


def polygon_centroid(x: ArrayLike, y: ArrayLike) -> tuple[float, float]:
    xs, ys = np.asarray(x, dtype=np.float64).ravel(), np.asarray(y, dtype=np.float64).ravel()
    xn, yn = np.roll(xs, -1), np.roll(ys, -1)
    cross = xs * yn - xn * ys
    area = 0.5 * cross.sum()
    if area == 0.0:
        raise ValueError("degenerate polygon (zero area)")
    return float(((xs + xn) * cross).sum() / (6.0 * area)), float(((ys + yn) * cross).sum() / (6.0 * area))


def mask_centroid(alpha: ArrayLike, beta: ArrayLike, captured: ArrayLike) -> tuple[float, float]:
    mask = np.asarray(captured, dtype=bool)
    if not np.any(mask):
        raise ValueError("the captured mask is empty: no shadow to extract")
    a, b = np.asarray(alpha, dtype=np.float64), np.asarray(beta, dtype=np.float64)
    return float(a[mask].mean()), float(b[mask].mean())


def _grid_geometry(alpha: Float, beta: Float) -> tuple[float, float, float]:
    """``(alpha_0, beta_0, d)`` of a regular camera grid (boundary.py:84-92)."""
    if alpha.ndim != 2 or alpha.shape != beta.shape or min(alpha.shape) < 2:
        raise ValueError("alpha and beta must be 2-D arrays of the same shape with >= 2 pixels per side")
    d = float(alpha[0, 1] - alpha[0, 0])
    if not d > 0.0:
        raise ValueError("alpha must increase along the columns")
    if not np.allclose(np.diff(alpha, axis=1), d) or not np.allclose(np.diff(beta, axis=0), -d):
        raise ValueError("the pixel grid must be regular with alpha increasing rightward and beta downward")
    return float(alpha[0, 0]), float(beta[0, 0]), d


def extract_boundary_from_mask(alpha: ArrayLike, beta: ArrayLike, captured: ArrayLike, *, n_angles: int = 360,
                               centre: tuple[float, float] | None = None, backend: Any = "hip") -> tuple[Float, Float]:
    """Sub-pixel edge radii ``r(phi_k)``, ``phi_k = 2 pi k / n_angles`` (boundary.py:116-170).

    The marching runs on the device (``kr_mask_edge_dev``, one wavefront per
    direction, 64 samples at a time); the checks, the captured centroid and
    the directions' cosines and sines are the host's, as the reference
    computes them.
    """
    import torch

    from kerrray_amd import _lib
    from kerrray_amd.backends import get_backend

    a = np.asarray(alpha, dtype=np.float64)
    b = np.asarray(beta, dtype=np.float64)
    mask = np.asarray(captured, dtype=bool)
    if mask.shape != a.shape:
        raise ValueError(f"captured has shape {mask.shape}, expected {a.shape}")
    if n_angles < 1:
        raise ValueError("n_angles must be >= 1")
    if mask[0, :].any() or mask[-1, :].any() or mask[:, 0].any() or mask[:, -1].any():
        raise ValueError("the captured region touches the image border; increase fov")
    a0, b0, d = _grid_geometry(a, b)
    ca, cb = mask_centroid(a, b, mask) if centre is None else (float(centre[0]), float(centre[1]))
    angles = 2.0 * np.pi * np.arange(n_angles, dtype=np.float64) / n_angles
    reach = math.hypot(a[0, -1] - a[0, 0], b[0, 0] - b[-1, 0]) + 2.0 * d
    engine = get_backend(backend) if isinstance(backend, str) else backend
    dev = torch.device("cuda", engine.handle.device)
    d_mask = torch.as_tensor(np.ascontiguousarray(mask.astype(np.uint8)), device=dev)
    d_cos = torch.as_tensor(np.cos(angles), device=dev)
    d_sin = torch.as_tensor(np.sin(angles), device=dev)
    radii = torch.empty(n_angles, dtype=torch.float64, device=dev)
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with engine.handle._lock:
        _lib.check(_lib.lib().kr_mask_edge_dev(engine.handle.ptr, d_mask.data_ptr(), mask.shape[0], mask.shape[1], a0,
                                               b0, d, ca, cb, d_cos.data_ptr(), d_sin.data_ptr(), n_angles, reach,
                                               radii.data_ptr(), flags.data_ptr(), ctypes.c_void_p(stream)),
                   "kr_mask_edge_dev")
    out = radii.cpu().numpy()
    _lib.raise_for_flags(int(flags.item()))
    return angles, out


def finite_distance_scale(st: Spacetime, radius: float) -> float:
    """``1 / sqrt(1 - 2M / r_o)``, which maps a ZAMO camera's image coordinates at ``r_o`` to
    Bardeen's: exact for ``a = 0``, to leading order otherwise (boundary.py:247-252)."""
    if not radius > 2.0 * st.mass:
        raise ValueError("the observer radius must exceed 2M for the finite-distance factor")
    return 1.0 / math.sqrt(1.0 - 2.0 * st.mass / radius)


def edge_error(radii: ArrayLike, angles: ArrayLike, centre: tuple[float, float],
               curve: tuple[ArrayLike, ArrayLike]) -> dict[str, float]:
    """``max``, ``rms`` and signed ``mean`` of traced radii minus the curve's radii (units of M)."""
    r = np.asarray(radii, dtype=np.float64)
    diff = r - polygon_radii(curve[0], curve[1], centre, angles)
    return {"max": float(np.max(np.abs(diff))), "rms": float(np.sqrt(np.mean(diff**2))),
            "mean": float(np.mean(diff))}
