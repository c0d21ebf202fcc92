"""Sub-pixel shadow edge of a captured mask (boundary.py:74-170).

The oracle's restatement (NumPy, ``oracle.mask_edge``) and the kernel source
built for the CPU (``mask_edge_radius``, one direction at a time) both equal
the reference's ``extract_boundary_from_mask`` bit for bit on four masks
(tests/golden/mask_edges.npz, written by tests/golden/make_golden.py): the
512 x 512 Kerr mask (a = 0.9, i = 60 deg), the two 64 x 64 ZAMO masks of docs/shadow.md
section 5 and a synthetic disk. The marching uses only IEEE arithmetic and
floor on values the host gives it (cos and sin of the directions), so the
device must match bit for bit too (tests/test_gpu_edge.py).
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np
import pytest
from conftest import GOLDEN, ROOT

CASES = ("c2", "zamo0", "zamo1", "disk")


@pytest.fixture(scope="module")
def gold() -> dict:
    with np.load(os.path.join(GOLDEN, "mask_edges.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case(gold, name):
    n, fov = int(gold[f"{name}_n"]), float(gold[f"{name}_fov"])
    c = fov * (2.0 * np.arange(n) + 1.0 - n) / n
    a, b = np.meshgrid(c, -c)
    mask = np.unpackbits(gold[f"{name}_mask"])[: n * n].reshape(n, n).astype(bool)
    return a, b, mask, gold[f"{name}_centre"], gold[f"{name}_angles"], gold[f"{name}_radii"]


def test_oracle_equals_the_reference(oracle, gold) -> None:
    for name in CASES:
        a, b, mask, centre, angles, radii = case(gold, name)
        ang, rad = oracle.mask_edge(a, b, mask, len(angles), None if name != "disk" else tuple(centre))
        assert np.array_equal(ang, angles) and np.array_equal(rad, radii), name


def test_kernel_source_equals_the_reference(gold) -> None:
    hostref = os.path.join(ROOT, "tests", "hostref")
    subprocess.run(["make", "-s", "-C", hostref], check=True)
    L = ctypes.CDLL(os.path.join(hostref, "libkr_hostref.so"))
    L.krh_mask_edge.restype = ctypes.c_int
    L.krh_mask_edge.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_double] * 5 + [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p]
    for name in CASES:
        a, b, mask, centre, angles, radii = case(gold, name)
        m = np.ascontiguousarray(mask.astype(np.uint8))
        d = float(a[0, 1] - a[0, 0])
        reach = math.hypot(a[0, -1] - a[0, 0], b[0, 0] - b[-1, 0]) + 2.0 * d
        cs, sn = np.cos(angles), np.sin(angles)
        out = np.empty(len(angles))
        bad = L.krh_mask_edge(m.ctypes.data, m.shape[0], m.shape[1], float(a[0, 0]), float(b[0, 0]), d,
                              float(centre[0]), float(centre[1]), cs.ctypes.data, sn.ctypes.data, len(angles), reach,
                              out.ctypes.data)
        assert bad == 0 and np.array_equal(out, radii), name
