"""Argument checks and limits of the image, disk and edge entry points (empty inputs, sizes, invalid options).

Every case here is turned into an error by the host before a launch, or is a
launch of zero rows; no kernel sees an invalid input.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
from conftest import options

from kerrray_amd import _lib
from kerrray_amd.camera import BardeenCamera, Camera, image_struct
from kerrray_amd.geometry import Spacetime

pytestmark = pytest.mark.gpu


def test_image_states_rejects_bad_reverse_and_sizes(hip) -> None:
    import torch

    L = _lib.lib()
    y = torch.empty((4, 8), dtype=torch.float64, device="cuda")
    st = Spacetime(1.0, 0.9)
    zamo = image_struct(st, Camera(radius=1000.0, inclination_deg=60.0, resolution=2))
    bardeen = image_struct(st, BardeenCamera(radius=1000.0, inclination_deg=60.0, resolution=2))
    assert L.kr_image_states_dev(hip.handle.ptr, ctypes.byref(zamo), 2, y.data_ptr(), None, None) == _lib.KR_ERR_INVALID
    assert L.kr_image_states_dev(hip.handle.ptr, ctypes.byref(bardeen), 0, y.data_ptr(), None, None) == _lib.KR_ERR_INVALID
    big = image_struct(st, Camera(radius=1000.0, inclination_deg=60.0, resolution=2))
    big.resolution = 46342
    assert L.kr_image_states_dev(hip.handle.ptr, ctypes.byref(big), 1, y.data_ptr(), None, None) == _lib.KR_ERR_INVALID
    bad_camera = image_struct(st, Camera(radius=1000.0, inclination_deg=60.0, resolution=2))
    bad_camera.camera = 7
    assert L.kr_image_states_dev(hip.handle.ptr, ctypes.byref(bad_camera), 1, y.data_ptr(), None, None) == _lib.KR_ERR_INVALID
    # The smallest image: 2 x 2 pixels.
    assert L.kr_image_states_dev(hip.handle.ptr, ctypes.byref(zamo), 1, y.data_ptr(), None, None) == _lib.KR_OK
    torch.cuda.synchronize()
    assert torch.isfinite(y).all()


def test_zamo_image_wider_than_the_sky_is_rejected_before_launch(hip) -> None:
    from kerrray_amd.shadow import compute_shadow

    integ, term = options(rtol=1e-6, atol=1e-8)
    with pytest.raises(ValueError, match="fov"):
        compute_shadow(Spacetime(1.0, 0.5), Camera(radius=10.0, inclination_deg=60.0, fov=9.0, resolution=8), integ,
                       term, backend=hip)


def test_empty_shading_and_edge_calls(hip) -> None:
    import torch

    L = _lib.lib()
    kd = _lib.KrDisk(6.0, 20.0, 3.0, 4, 1)
    z = torch.empty(1, dtype=torch.float64, device="cuda")
    assert L.kr_disk_shade_dev(hip.handle.ptr, ctypes.byref(kd), 1.0, 0.5, _lib.KR_FLOAT64, None, None, 0, None, None,
                               None, None, None) == _lib.KR_OK
    mask = torch.zeros((4, 4), dtype=torch.uint8, device="cuda")
    assert L.kr_mask_edge_dev(hip.handle.ptr, mask.data_ptr(), 4, 4, -1.5, 1.5, 1.0, 0.0, 0.0, None, None, 0, 6.0,
                              None, None, None) == _lib.KR_OK
    bad = _lib.KrDisk(6.0, 20.0, 3.0, 5, 1)
    assert L.kr_disk_shade_dev(hip.handle.ptr, ctypes.byref(bad), 1.0, 0.5, _lib.KR_FLOAT64, z.data_ptr(),
                               z.data_ptr(), 1, z.data_ptr(), z.data_ptr(), z.data_ptr(), None, None) == _lib.KR_ERR_INVALID
    assert L.kr_mask_edge_dev(hip.handle.ptr, mask.data_ptr(), 1, 4, -1.5, 1.5, 1.0, 0.0, 0.0, z.data_ptr(),
                              z.data_ptr(), 1, 6.0, z.data_ptr(), None, None) == _lib.KR_ERR_INVALID


def test_render_disk_needs_matching_event_constants(hip) -> None:
    from kerrray_amd.backends import HostBatch, kernel_params
    from kerrray_amd.renderer import DiskModel, event_options

    st = Spacetime(1.0, 0.5)
    cam = Camera(radius=1000.0, inclination_deg=75.0, fov=20.0, resolution=4)
    disk = DiskModel(r_in=6.0, r_out=20.0)
    integ, term = options(rtol=1e-6, atol=1e-8)
    other = DiskModel(r_in=7.0, r_out=20.0)
    p = kernel_params(st, integ, term, event_options(other))
    out = HostBatch(cam.n_rays, np.dtype(np.float64), True)
    arrays = [np.empty(cam.n_rays) for _ in range(3)]
    kd = disk.kr()
    status = _lib.lib().kr_render_disk(hip.handle.ptr, ctypes.byref(p), ctypes.byref(image_struct(st, cam)),
                                       ctypes.byref(kd), ctypes.byref(out.outputs), *(a.ctypes.data for a in arrays),
                                       None)
    assert status == _lib.KR_ERR_INVALID and "r_in" in _lib.last_error()


@pytest.mark.parametrize("missing", ["lam", "max_carter_drift", "n_rejected"])
def test_planned_device_trace_needs_every_output_array(hip, missing) -> None:
    """The planned kernels store every BatchResult array of a row, so each must be given."""
    import torch

    from bench import device_outputs
    from kerrray_amd.backends import kernel_params
    from kerrray_amd.planned import PlannedOptions, edge_table
    from kerrray_amd.shadow import kr_image

    st = Spacetime(1.0, 0.9)
    integ, term = options(rtol=1e-6, atol=1e-8)
    n = 16 * 16
    bufs, outs = device_outputs(torch, n, torch.float64)
    setattr(outs, missing, None)
    table = edge_table(st, 60.0)
    radii = torch.from_numpy(table.radii).cuda()
    L = _lib.lib()
    nbytes = L.kr_planned_workspace_bytes(n, table.n_angles)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    p, img, ko = kernel_params(st, integ, term), kr_image(st, 1000.0, 60.0, 10.0, 16), PlannedOptions().kr()
    status = L.kr_trace_image_planned_dev(hip.handle.ptr, ctypes.byref(p), ctypes.byref(img), ctypes.byref(ko),
                                          table.centre[0], table.centre[1], radii.data_ptr(), table.n_angles,
                                          ctypes.byref(outs), work.data_ptr(), nbytes, None)
    assert status == _lib.KR_ERR_INVALID and "output" in _lib.last_error()
