"""The precision form's verify/repair reaches the FP64 mask as a fixed point.

tests/hostref compiles kerrray_amd/csrc/kr_repair.h (the helpers the GPU's
verify and repair kernels are made of) for the CPU and runs the rounds
serially, with the FP64 trace of a pixel replaced by a planted class. The
planted FP32 masks misclassify more pixels than one round can reach: each
round only re-traces pixels within band_px of a class change, so a wrong
ring several band widths wide needs several rounds, and a fixed two-round
scheme would leave FP32 classes behind.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

HOSTREF = os.path.join(ROOT, "tests", "hostref")
CAPTURED, ESCAPED, FAILURE = 2, 1, 4
MARK_FP64 = 1
REPAIR_ROUNDS = 4  # KR_REPAIR_ROUNDS of include/kerrray_hip.h
FLAG_REPAIR = 32


@pytest.fixture(scope="module")
def hostref():
    subprocess.run(["make", "-s", "-C", HOSTREF], check=True, capture_output=True)
    L = ctypes.CDLL(os.path.join(HOSTREF, "libkr_hostref.so"))
    L.krh_repair_rounds.restype = ctypes.c_int
    L.krh_repair_rounds.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p]
    return L


def disk(n: int, radius: float, centre: float | None = None) -> np.ndarray:
    c = (n - 1) / 2 if centre is None else centre
    i, j = np.mgrid[0:n, 0:n]
    return np.where((i - c) ** 2 + (j - c) ** 2 <= radius**2, CAPTURED, ESCAPED).astype(np.int64).ravel()


def run(L, n, band, traced, fp64, marks, rounds):
    state = traced.copy()
    marks = marks.copy()
    listed = np.zeros(rounds + 1, np.int64)
    flags = L.krh_repair_rounds(n, band, state.ctypes.data, fp64.ctypes.data, marks.ctypes.data, rounds,
                                listed.ctypes.data)
    return state, marks, listed, flags


def unverified(state: np.ndarray, marks: np.ndarray, n: int, band: int) -> int:
    """Pixels not traced in FP64 that the verify rule would still send to FP64 (a brute-force restatement)."""
    s = state.reshape(n, n)
    cap = s == CAPTURED
    m = marks.reshape(n, n) & MARK_FP64
    bad = 0
    for i in range(n):
        for j in range(n):
            if m[i, j]:
                continue
            if s[i, j] not in (CAPTURED, ESCAPED):
                bad += 1
                continue
            win = cap[max(0, i - band):i + band + 1, max(0, j - band):j + band + 1]
            bad += int(np.any(win != cap[i, j]))
    return bad


@pytest.mark.parametrize("band", [1, 2])
def test_wide_misclassified_ring_needs_several_rounds(hostref, band):
    n = 48
    fp64 = disk(n, 14.0)
    traced = disk(n, 6.0)  # FP32 classes: an 8-pixel ring wrongly escaped
    marks = np.zeros(n * n, np.uint8)
    state, out_marks, listed, flags = run(hostref, n, band, traced, fp64, marks, REPAIR_ROUNDS + 8)
    assert flags == 0
    used = int(np.count_nonzero(listed[:-1]))
    assert used >= 3, listed  # one round, or the former two, would not reach the true edge
    assert np.array_equal(state, fp64)
    assert unverified(state, out_marks, n, band) == 0
    # Every pixel listed once: the lists sum to the pixels that ended marked FP64.
    assert listed[:-1].sum() == np.count_nonzero(out_marks & MARK_FP64)


def test_round_limit_raises_the_flag(hostref):
    n, band = 48, 2
    fp64 = disk(n, 14.0)
    traced = disk(n, 6.0)
    marks = np.zeros(n * n, np.uint8)
    state, _, listed, flags = run(hostref, n, band, traced, fp64, marks, 2)
    assert flags == FLAG_REPAIR and listed[2] > 0
    assert not np.array_equal(state, fp64)  # the FP32 classes left behind are what the flag reports


def test_one_round_suffices_when_nothing_changes_class(hostref):
    n, band = 40, 2
    fp64 = disk(n, 10.0)
    # A pixel that failed in FP32 and is escaped in FP64, far from the edge.
    traced = fp64.copy()
    traced[3 * n + 3] = FAILURE
    marks = np.zeros(n * n, np.uint8)
    state, out_marks, listed, flags = run(hostref, n, band, traced, fp64, marks, REPAIR_ROUNDS)
    assert flags == 0 and listed[1:].sum() == 0
    assert np.array_equal(state, fp64)
    assert unverified(state, out_marks, n, band) == 0


def test_pixels_traced_in_fp64_are_never_relisted(hostref):
    n, band = 48, 2
    fp64 = disk(n, 14.0)
    traced = disk(n, 6.0)
    marks = np.zeros(n * n, np.uint8)
    # The planned FP64 band around the true edge: already exact.
    i, j = np.mgrid[0:n, 0:n]
    rho = np.hypot(i - (n - 1) / 2, j - (n - 1) / 2).ravel()
    band_pixels = np.abs(rho - 14.0) <= 3.0
    marks[band_pixels] = MARK_FP64
    traced[band_pixels] = fp64[band_pixels]
    state, out_marks, listed, flags = run(hostref, n, band, traced, fp64, marks, REPAIR_ROUNDS + 8)
    assert flags == 0
    assert np.array_equal(state, fp64)
    assert unverified(state, out_marks, n, band) == 0
    assert listed[:-1].sum() == np.count_nonzero(out_marks & MARK_FP64) - np.count_nonzero(band_pixels)
