"""Shadow-edge geometry on the image plane (host side).

Polygon radii and centroid, the sub-pixel edge of a captured mask and the
comparison with an analytic curve, as in src/kerrray/raytracing/boundary.py:74-307:

* ``polygon_radii``: exact ray/segment intersection, smallest t > 0 with
  0 <= s <= 1 (widened by 1e-9 so a ray through a vertex is not lost);
* ``polygon_centroid``: shoelace area centroid;
* ``extract_boundary_from_mask``: bilinear 0.5 level set of the mask (zero
  outside the image) marched along polar rays in steps of 1/8 pixel;
* ``edge_error``: radial differences between traced edge radii and the curve.
"""

from __future__ import annotations

import math

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


def polygon_radii(alpha: ArrayLike, beta: ArrayLike, centre: tuple[float, float], angles: ArrayLike) -> Float:
    """Distance from ``centre`` to the closed polygon along each angle."""
    px = np.asarray(alpha, dtype=np.float64).ravel()
    py = np.asarray(beta, dtype=np.float64).ravel()
    if px.size < 3 or px.shape != py.shape:
        raise ValueError("a polygon needs at least three (alpha, beta) points")
    phi = np.asarray(angles, dtype=np.float64).ravel()
    qx, qy = px - float(centre[0]), py - float(centre[1])
    ex, ey = np.roll(qx, -1) - qx, np.roll(qy, -1) - qy
    dx, dy = np.cos(phi)[:, None], np.sin(phi)[:, None]
    det = dx * (-ey[None, :]) - dy * (-ex[None, :])
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (qx[None, :] * (-ey[None, :]) - qy[None, :] * (-ex[None, :])) / det
        s = (dx * qy[None, :] - dy * qx[None, :]) / det
    valid = np.isfinite(t) & (t > 0.0) & (s >= -EDGE_TOLERANCE) & (s <= 1.0 + EDGE_TOLERANCE)
    radii = np.where(valid, t, np.inf).min(axis=1)
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


def _bilinear(field: Float, u: Float, v: Float) -> Float:
    n_rows, n_cols = field.shape
    padded = np.pad(field, 1)
    up, vp = u + 1.0, v + 1.0
    inside = (up >= 0.0) & (up <= n_cols) & (vp >= 0.0) & (vp <= n_rows)
    j0 = np.clip(np.floor(up).astype(np.intp), 0, n_cols)
    i0 = np.clip(np.floor(vp).astype(np.intp), 0, n_rows)
    fu = np.clip(up - j0, 0.0, 1.0)
    fv = np.clip(vp - i0, 0.0, 1.0)
    j1 = np.minimum(j0 + 1, n_cols + 1)
    i1 = np.minimum(i0 + 1, n_rows + 1)
    value = (padded[i0, j0] * (1.0 - fu) * (1.0 - fv) + padded[i0, j1] * fu * (1.0 - fv)
             + padded[i1, j0] * (1.0 - fu) * fv + padded[i1, j1] * fu * fv)
    return np.where(inside, value, 0.0)


def extract_boundary_from_mask(alpha: ArrayLike, beta: ArrayLike, captured: ArrayLike, *, n_angles: int = 360,
                               centre: tuple[float, float] | None = None) -> tuple[Float, Float]:
    """Sub-pixel edge radii ``r(phi_k)``, ``phi_k = 2 pi k / n_angles`` (boundary.py:116-170)."""
    a = np.asarray(alpha, dtype=np.float64)
    b = np.asarray(beta, dtype=np.float64)
    mask = np.asarray(captured, dtype=bool)
    if mask.shape != a.shape:
        raise ValueError(f"captured has shape {mask.shape}, expected {a.shape}")
    if mask[0, :].any() or mask[-1, :].any() or mask[:, 0].any() or mask[:, -1].any():
        raise ValueError("the captured region touches the image border; increase fov")
    d = float(a[0, 1] - a[0, 0])
    a0, b0 = float(a[0, 0]), float(b[0, 0])
    ca, cb = mask_centroid(a, b, mask) if centre is None else (float(centre[0]), float(centre[1]))
    field = mask.astype(np.float64)
    angles = 2.0 * np.pi * np.arange(n_angles, dtype=np.float64) / n_angles
    ds = STEP_FRACTION * d
    reach = math.hypot(a[0, -1] - a[0, 0], b[0, 0] - b[-1, 0]) + 2.0 * d
    s = ds * np.arange(int(math.ceil(reach / ds)) + 1, dtype=np.float64)
    x = ca + np.cos(angles)[:, None] * s[None, :]
    y = cb + np.sin(angles)[:, None] * s[None, :]
    f = _bilinear(field, (x - a0) / d, (b0 - y) / d)
    if np.any(f[:, 0] < 0.5):
        raise ValueError("the centre lies outside the captured region")
    k = np.argmax(f < 0.5, axis=1)
    rows = np.arange(n_angles)
    f_in, f_out = f[rows, k - 1], f[rows, k]
    return angles, s[k - 1] + ds * (f_in - 0.5) / (f_in - f_out)


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
