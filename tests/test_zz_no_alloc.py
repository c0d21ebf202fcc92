"""The allocation-free planned trace equals the workspace entry point's, call after call.

kr_trace_image_planned_noalloc runs the planned kernels on the library's
static device scratch; kr_trace_image_planned_dev runs them on a caller
workspace. Two calls in a row, the second on a smaller image, must give
every output array bit for bit as the workspace entry point gives it; an
image above the scratch's limit is refused before any launch.
"""

from __future__ import annotations

import ctypes

import pytest
from conftest import options

from kerrray_amd import _lib
from kerrray_amd.geometry import Spacetime

pytestmark = pytest.mark.gpu

FIELDS = ("Y", "lam", "n_steps", "n_rejected", "state", "m0", "m1", "m2", "m3")


def _call(hip, spin, incl, res, rtol, fp32_bulk):
    import torch

    from bench import device_outputs
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.planned import PlannedOptions, edge_table
    from kerrray_amd.shadow import kr_image

    L = _lib.lib()
    st = Spacetime(1.0, spin)
    integ, term = options(rtol=rtol, atol=1e-2 * rtol)
    n = res * res
    table = edge_table(st, incl)
    radii = torch.from_numpy(table.radii).cuda()
    p, img = kernel_params(st, integ, term), kr_image(st, 1000.0, incl, 10.0, res)
    ko = PlannedOptions(fp32_bulk=fp32_bulk).kr()
    stream = torch.cuda.current_stream()
    ref_bufs, ref_outs = device_outputs(torch, n, torch.float64)
    nbytes = L.kr_planned_workspace_bytes(n, table.n_angles)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    _lib.check(L.kr_trace_image_planned_dev(hip.handle.ptr, ctypes.byref(p), ctypes.byref(img), ctypes.byref(ko),
                                            table.centre[0], table.centre[1], radii.data_ptr(), table.n_angles,
                                            ctypes.byref(ref_outs), work.data_ptr(), nbytes,
                                            ctypes.c_void_p(stream.cuda_stream)), "kr_trace_image_planned_dev")
    bufs, outs = device_outputs(torch, n, torch.float64)
    flags = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(L.kr_trace_image_planned_noalloc(hip.handle.ptr, ctypes.byref(p), ctypes.byref(img),
                                                ctypes.byref(ko), table.centre[0], table.centre[1],
                                                radii.data_ptr(), table.n_angles, ctypes.byref(outs),
                                                flags.data_ptr(), ctypes.c_void_p(stream.cuda_stream)),
               "kr_trace_image_planned_noalloc")
    torch.cuda.synchronize()
    return ref_bufs, bufs, int(flags.item())


def test_noalloc_equals_the_workspace_entry_point_twice(hip) -> None:
    import torch

    for spin, incl, res, rtol, fp32 in [(0.9, 60.0, 512, 1e-6, True), (0.5, 17.0, 128, 1e-7, False)]:
        print(f"[noalloc] a={spin} i={incl} {res}^2 rtol {rtol} fp32_bulk={fp32}", flush=True)
        ref, got, flags = _call(hip, spin, incl, res, rtol, fp32)
        assert flags == 0
        for name in FIELDS:
            assert torch.equal(ref[name], got[name]), name
        assert int((got["state"] == 2).sum()) > 0


def test_noalloc_refuses_images_beyond_its_scratch(hip) -> None:
    import torch

    from bench import device_outputs
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.planned import PlannedOptions, edge_table
    from kerrray_amd.shadow import kr_image

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-6, atol=1e-8)
    table = edge_table(st, 60.0)
    radii = torch.from_numpy(table.radii).cuda()
    bufs, outs = device_outputs(torch, 1, torch.float64)  # never written: the call is refused first
    p, img, ko = kernel_params(st, integ, term), kr_image(st, 1000.0, 60.0, 10.0, 1025), PlannedOptions().kr()
    status = _lib.lib().kr_trace_image_planned_noalloc(
        hip.handle.ptr, ctypes.byref(p), ctypes.byref(img), ctypes.byref(ko), table.centre[0], table.centre[1],
        radii.data_ptr(), table.n_angles, ctypes.byref(outs), None, None)
    assert status == _lib.KR_ERR_INVALID and "2^20" in _lib.last_error()
    small = kr_image(st, 1000.0, 60.0, 10.0, 16)
    outs.lam = None
    status = _lib.lib().kr_trace_image_planned_noalloc(
        hip.handle.ptr, ctypes.byref(p), ctypes.byref(small), ctypes.byref(ko), table.centre[0], table.centre[1],
        radii.data_ptr(), table.n_angles, ctypes.byref(outs), None, None)
    assert status == _lib.KR_ERR_INVALID and "output" in _lib.last_error()
