"""Mask edge extraction on the device equals the reference bit for bit (tests/golden/mask_edges.npz)."""

from __future__ import annotations

import os

import numpy as np
import pytest
from conftest import GOLDEN
from test_mask_edge import CASES, case

from kerrray_amd.boundary import extract_boundary_from_mask

pytestmark = pytest.mark.gpu


def test_device_edge_equals_the_reference(hip) -> None:
    with np.load(os.path.join(GOLDEN, "mask_edges.npz"), allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    for name in CASES:
        a, b, mask, centre, angles, radii = case(gold, name)
        ang, rad = extract_boundary_from_mask(a, b, mask, n_angles=len(angles),
                                              centre=None if name != "disk" else tuple(centre), backend=hip)
        print(f"[mask edge {name}] {mask.shape[0]}^2, {len(angles)} directions: bitwise {np.mean(rad == radii):.4f}",
              flush=True)
        assert np.array_equal(ang, angles) and np.array_equal(rad, radii), name
    a, b, mask, *_ = case(gold, "disk")
    with pytest.raises(ValueError):
        extract_boundary_from_mask(a, b, mask, n_angles=8, centre=(8.0, 8.0), backend=hip)
