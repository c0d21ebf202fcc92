"""ctypes binding of ``libkrhip.so`` (declarations in ``include/kerrray_hip.h``).

The library is built in-tree by :mod:`kerrray_amd.build`. Loading it never
falls back to anything else: a missing or unloadable library raises
:class:`LibraryMissing`, which the backend registry reports as the ``hip``
backend being unavailable.
"""

from __future__ import annotations

import ctypes
import os
import threading

from kerrray_amd.build import LIB_PATH

KR_OK, KR_ERR_INVALID, KR_ERR_DEVICE, KR_ERR_RADICAND, KR_ERR_UNSUPPORTED, KR_ERR_NOT_NULL, KR_ERR_REPAIR = range(7)
KR_FLOAT64, KR_FLOAT32 = 0, 1
KR_FLAG_RADICAND, KR_FLAG_NOT_NULL, KR_FLAG_SKY, KR_FLAG_DISK, KR_FLAG_EDGE, KR_FLAG_REPAIR = 1, 2, 4, 8, 16, 32
KR_REPAIR_ROUNDS = 4
KR_NOALLOC_MAX_PIXELS, KR_NOALLOC_MAX_ANGLES = 1 << 20, 8192
KR_CAMERA_BARDEEN, KR_CAMERA_ZAMO = 0, 1
ABI_VERSION = 7

# Every symbol the header declares (checked by the tests).
EXPORTED = (
    "kr_abi_version",
    "kr_device_count",
    "kr_create",
    "kr_destroy",
    "kr_last_error",
    "kr_integrate_batch",
    "kr_integrate_batch_dev",
    "kr_trace_image",
    "kr_trace_image_dev",
    "kr_trace_points_dev",
    "kr_image_states_dev",
    "kr_point_states_dev",
    "kr_zamo_frame",
    "kr_disk_shade_dev",
    "kr_mask_edge_dev",
    "kr_render_disk",
    "kr_render_disk_dev",
    "kr_bisect_edge",
    "kr_impact_search",
    "kr_batch_create",
    "kr_batch_photons",
    "kr_batch_run",
    "kr_batch_graph_launch",
    "kr_batch_flags",
    "kr_batch_destroy",
    "kr_batch_create_precision",
    "kr_batch_stats",
    "kr_planned_workspace_bytes",
    "kr_trace_image_planned",
    "kr_trace_image_planned_dev",
    "kr_planned_read_stats",
    "kr_trace_image_planned_noalloc",
    "kr_tape_create",
    "kr_tape_recorded",
    "kr_tape_steps",
    "kr_tape_record_ms",
    "kr_tape_jacobian_ms",
    "kr_tape_vjp",
    "kr_tape_flags",
    "kr_tape_destroy",
    "kr_escape_directions_dev",
)


class LibraryMissing(RuntimeError):
    """``libkrhip.so`` is absent or cannot be loaded."""


class KrError(RuntimeError):
    """A non-OK status from the C ABI."""

    def __init__(self, status: int, message: str) -> None:
        super().__init__(message)
        self.status = status


c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)


class KrParams(ctypes.Structure):
    _fields_ = [
        ("geom", ctypes.c_double * 7),
        ("consts", ctypes.c_double * 21),
        ("params", ctypes.c_double * 3),
        ("ipar", ctypes.c_int64 * 4),
        ("tab_a", ctypes.c_double * 49),
        ("tab_e", ctypes.c_double * 7),
        ("rk4_b", ctypes.c_double * 4),
        ("rk4_c", ctypes.c_double * 4),
        ("dtype", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class KrOutputs(ctypes.Structure):
    _fields_ = [
        ("Y", ctypes.c_void_p),
        ("lam", ctypes.c_void_p),
        ("n_steps", ctypes.c_void_p),
        ("n_rejected", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
        ("max_null_error", ctypes.c_void_p),
        ("max_energy_drift", ctypes.c_void_p),
        ("max_lz_drift", ctypes.c_void_p),
        ("max_carter_drift", ctypes.c_void_p),
        ("event_Y", ctypes.c_void_p),
        ("event_hit", ctypes.c_void_p),
    ]


class KrImage(ctypes.Structure):
    _fields_ = [
        ("mass", ctypes.c_double),
        ("spin", ctypes.c_double),
        ("r_o", ctypes.c_double),
        ("theta_o", ctypes.c_double),
        ("fov", ctypes.c_double),
        ("resolution", ctypes.c_int64),
        ("camera", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("phi_o", ctypes.c_double),
    ]


class KrDisk(ctypes.Structure):
    _fields_ = [
        ("r_in", ctypes.c_double),
        ("r_out", ctypes.c_double),
        ("emissivity_index", ctypes.c_double),
        ("redshift_power", ctypes.c_int32),
        ("prograde", ctypes.c_int32),
    ]


class KrJob(ctypes.Structure):
    _fields_ = [("params", KrParams), ("image", KrImage)]


class KrPlannedOpts(ctypes.Structure):
    _fields_ = [
        ("horizon_margin", ctypes.c_double),
        ("axis_sin", ctypes.c_double),
        ("turn_rmax", ctypes.c_double),
        ("turn_drdl", ctypes.c_double),
        ("band_px", ctypes.c_int32),
        ("fp32_bulk", ctypes.c_int32),
        ("refill_interval", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("longest_first", ctypes.c_int32),
        ("band_priority", ctypes.c_int32),
        ("waves_per_simd", ctypes.c_int32),
        ("split_roles", ctypes.c_int32),
        ("axis_lz", ctypes.c_double),
        ("grid_fraction", ctypes.c_double),
    ]


class KrPlannedStats(ctypes.Structure):
    _fields_ = [
        ("planned_band", ctypes.c_int64),
        ("planned_inside", ctypes.c_int64),
        ("planned_outside", ctypes.c_int64),
        ("repaired", ctypes.c_int64),
        ("fp64_attempts", ctypes.c_int64),
        ("repair_rounds", ctypes.c_int64),
        ("unrepaired", ctypes.c_int64),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class KrImpactOpts(ctypes.Structure):
    _fields_ = [
        ("r0", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("cos2_theta", ctypes.c_double),
        ("sin2_theta", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("capture_radius", ctypes.c_double),
        ("max_iterations", ctypes.c_int64),
        ("n_probe", ctypes.c_int32),
        ("horizon_crossing_as_capture", ctypes.c_int32),
    ]


class KrImpactResult(ctypes.Structure):
    _fields_ = [
        ("rays", ctypes.c_int64),
        ("max_steps", ctypes.c_int64),
        ("crossings", ctypes.c_int64),
        ("error", ctypes.c_int32),
        ("error_round", ctypes.c_int32),
    ]


class KrBatchPrecision(ctypes.Structure):
    _fields_ = [
        ("band_px", ctypes.c_int32),
        ("n_angles", ctypes.c_int32),
        ("axis_lz", ctypes.c_double),
        ("fp32_min_rtol", ctypes.c_double),
        ("centre_alpha", ctypes.c_void_p),
        ("centre_beta", ctypes.c_void_p),
        ("radii", ctypes.c_void_p),
    ]


class KrTune(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int32),
        ("waves_per_simd", ctypes.c_int32),
        ("refill_interval", ctypes.c_int32),
        ("order", ctypes.c_int32),
    ]


_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def _declare(L: ctypes.CDLL) -> None:
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    L.kr_abi_version.restype = ctypes.c_int
    L.kr_abi_version.argtypes = []
    L.kr_device_count.restype = ctypes.c_int
    L.kr_device_count.argtypes = []
    L.kr_create.restype = ctypes.c_int
    L.kr_create.argtypes = [ctypes.c_int, P(vp)]
    L.kr_destroy.restype = ctypes.c_int
    L.kr_destroy.argtypes = [vp]
    L.kr_last_error.restype = ctypes.c_char_p
    L.kr_last_error.argtypes = []
    L.kr_integrate_batch.restype = ctypes.c_int
    L.kr_integrate_batch.argtypes = [vp, P(KrParams), vp, ctypes.c_int64, P(KrOutputs), c_double_p]
    L.kr_integrate_batch_dev.restype = ctypes.c_int
    L.kr_integrate_batch_dev.argtypes = [vp, P(KrParams), vp, ctypes.c_int64, P(KrOutputs), vp]
    L.kr_trace_image.restype = ctypes.c_int
    L.kr_trace_image.argtypes = [vp, P(KrParams), P(KrImage), P(KrOutputs), c_double_p]
    L.kr_trace_image_dev.restype = ctypes.c_int
    L.kr_trace_image_dev.argtypes = [vp, P(KrParams), P(KrImage), P(KrOutputs), vp, vp]
    L.kr_trace_points_dev.restype = ctypes.c_int
    L.kr_trace_points_dev.argtypes = [vp, P(KrParams), P(KrImage), vp, vp, ctypes.c_int64, P(KrOutputs), vp, vp]
    L.kr_bisect_edge.restype = ctypes.c_int
    L.kr_bisect_edge.argtypes = [
        vp, P(KrParams), P(KrImage), ctypes.c_double, ctypes.c_double, vp, vp, ctypes.c_int64, vp, vp,
        ctypes.c_int32, ctypes.c_int32, c_int64_p, c_int64_p, c_double_p,
    ]
    L.kr_impact_search.restype = ctypes.c_int
    L.kr_impact_search.argtypes = [vp, P(KrParams), P(KrImpactOpts), ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp,
                                   P(KrImpactResult), c_double_p]
    L.kr_batch_create.restype = ctypes.c_int
    L.kr_batch_create.argtypes = [vp, P(KrJob), ctypes.c_int32, P(KrTune), P(KrOutputs), P(vp)]
    L.kr_batch_photons.restype = ctypes.c_int64
    L.kr_batch_photons.argtypes = [vp]
    L.kr_batch_run.restype = ctypes.c_int
    L.kr_batch_run.argtypes = [vp, vp]
    L.kr_batch_graph_launch.restype = ctypes.c_int
    L.kr_batch_graph_launch.argtypes = [vp, vp]
    L.kr_batch_flags.restype = ctypes.c_int
    L.kr_batch_flags.argtypes = [vp, P(ctypes.c_int32)]
    L.kr_batch_destroy.restype = ctypes.c_int
    L.kr_batch_destroy.argtypes = [vp]
    L.kr_batch_create_precision.restype = ctypes.c_int
    L.kr_batch_create_precision.argtypes = [vp, P(KrJob), ctypes.c_int32, P(KrTune), P(KrBatchPrecision), P(KrOutputs),
                                            P(vp)]
    L.kr_batch_stats.restype = ctypes.c_int
    L.kr_batch_stats.argtypes = [vp, P(KrPlannedStats)]
    L.kr_planned_workspace_bytes.restype = ctypes.c_size_t
    L.kr_planned_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32]
    L.kr_trace_image_planned.restype = ctypes.c_int
    L.kr_trace_image_planned.argtypes = [vp, P(KrParams), P(KrImage), P(KrPlannedOpts), ctypes.c_double,
                                         ctypes.c_double, vp, ctypes.c_int32, P(KrOutputs), P(KrPlannedStats),
                                         c_double_p]
    L.kr_trace_image_planned_dev.restype = ctypes.c_int
    L.kr_trace_image_planned_dev.argtypes = [vp, P(KrParams), P(KrImage), P(KrPlannedOpts), ctypes.c_double,
                                             ctypes.c_double, vp, ctypes.c_int32, P(KrOutputs), vp, ctypes.c_size_t,
                                             vp]
    L.kr_planned_read_stats.restype = ctypes.c_int
    L.kr_planned_read_stats.argtypes = [vp, ctypes.c_size_t, P(KrPlannedStats), vp]
    L.kr_trace_image_planned_noalloc.restype = ctypes.c_int
    L.kr_trace_image_planned_noalloc.argtypes = [vp, P(KrParams), P(KrImage), P(KrPlannedOpts), ctypes.c_double,
                                                 ctypes.c_double, vp, ctypes.c_int32, P(KrOutputs), vp, vp]
    L.kr_tape_create.restype = ctypes.c_int
    L.kr_tape_create.argtypes = [vp, P(KrParams), P(KrImage), vp, vp, ctypes.c_int64, ctypes.c_int32, P(KrOutputs),
                                 vp, vp, P(vp)]
    L.kr_tape_recorded.restype = ctypes.c_int64
    L.kr_tape_recorded.argtypes = [vp]
    L.kr_tape_steps.restype = ctypes.c_int64
    L.kr_tape_steps.argtypes = [vp]
    L.kr_tape_record_ms.restype = ctypes.c_double
    L.kr_tape_record_ms.argtypes = [vp]
    L.kr_tape_jacobian_ms.restype = ctypes.c_double
    L.kr_tape_jacobian_ms.argtypes = [vp]
    L.kr_tape_vjp.restype = ctypes.c_int
    L.kr_tape_vjp.argtypes = [vp, vp, vp, vp, vp]
    L.kr_tape_flags.restype = ctypes.c_int
    L.kr_tape_flags.argtypes = [vp, P(ctypes.c_int32)]
    L.kr_tape_destroy.restype = ctypes.c_int
    L.kr_tape_destroy.argtypes = [vp]
    L.kr_image_states_dev.restype = ctypes.c_int
    L.kr_image_states_dev.argtypes = [vp, P(KrImage), ctypes.c_int32, vp, vp, vp]
    L.kr_point_states_dev.restype = ctypes.c_int
    L.kr_point_states_dev.argtypes = [vp, P(KrImage), ctypes.c_int32, vp, vp, ctypes.c_int64, vp, vp, vp]
    L.kr_disk_shade_dev.restype = ctypes.c_int
    L.kr_disk_shade_dev.argtypes = [vp, P(KrDisk), ctypes.c_double, ctypes.c_double, ctypes.c_int32, vp, vp,
                                    ctypes.c_int64, vp, vp, vp, vp, vp]
    L.kr_mask_edge_dev.restype = ctypes.c_int
    L.kr_mask_edge_dev.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_double] * 5 + [
        vp, vp, ctypes.c_int64, ctypes.c_double, vp, vp, vp]
    L.kr_render_disk.restype = ctypes.c_int
    L.kr_render_disk.argtypes = [vp, P(KrParams), P(KrImage), P(KrDisk), P(KrOutputs), vp, vp, vp, c_double_p]
    L.kr_render_disk_dev.restype = ctypes.c_int
    L.kr_render_disk_dev.argtypes = [vp, P(KrParams), P(KrImage), P(KrDisk), P(KrOutputs), vp, vp, vp, vp, vp]
    L.kr_zamo_frame.restype = ctypes.c_int
    L.kr_zamo_frame.argtypes = [P(KrImage), c_double_p, c_double_p, c_double_p]
    L.kr_escape_directions_dev.restype = ctypes.c_int
    L.kr_escape_directions_dev.argtypes = [vp, P(KrParams), vp, vp, ctypes.c_int64, vp, vp, vp]


def _share_torch_runtime() -> None:
    """Load PyTorch's HIP runtime first, when PyTorch is installed.

    PyTorch-ROCm ships its own ``libamdhip64.so`` with the soname
    ``libamdhip64.so.7``, the library this one needs. Loaded first, it is the
    one the dynamic linker then binds ``libkrhip.so`` to, so the process has
    a single HIP runtime and PyTorch's streams, events and allocations are
    valid arguments of the ``_dev`` entry points. Loaded the other way round,
    two runtimes would coexist and PyTorch would not see the GPU.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> ctypes.CDLL:
    """Load ``libkrhip.so`` once (raises :class:`LibraryMissing`)."""
    global _lib
    with _lock:
        if _lib is None:
            _share_torch_runtime()
            if not os.path.exists(LIB_PATH):
                raise LibraryMissing(
                    f"{LIB_PATH} is not built; run `python -m kerrray_amd.build` (needs ROCm hipcc)"
                )
            try:
                L = ctypes.CDLL(LIB_PATH)
            except OSError as exc:
                raise LibraryMissing(f"cannot load {LIB_PATH}: {exc}") from exc
            _declare(L)
            if L.kr_abi_version() != ABI_VERSION:
                raise LibraryMissing(f"{LIB_PATH} has ABI {L.kr_abi_version()}, expected {ABI_VERSION}")
            _lib = L
        return _lib


def last_error() -> str:
    msg = lib().kr_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(status: int, what: str) -> None:
    """Raise the Python exception the reference raises for this kind of failure.

    Invalid arguments and initial conditions raise ``ValueError``; the rest,
    including a ZAMO photon that is not null (``RuntimeError`` in the
    reference, camera.py:290-291,366-371), raise :class:`KrError`, a
    ``RuntimeError``.
    """
    if status == KR_OK:
        return
    msg = f"{what}: {last_error()}"
    if status in (KR_ERR_INVALID, KR_ERR_RADICAND, KR_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise KrError(status, msg)


def raise_for_flags(flags: int) -> None:
    """Raise what a device flag word (``KR_FLAG_*`` bits) stands for, as :func:`check` would."""
    if flags & KR_FLAG_REPAIR:
        raise KrError(KR_ERR_REPAIR, "the FP64 verify/repair of the mask did not reach a fixed point within "
                                     f"{KR_REPAIR_ROUNDS} rounds")
    if flags & KR_FLAG_RADICAND:
        raise ValueError("a Carter radicand is negative beyond round-off at the observer")
    if flags & KR_FLAG_SKY:
        raise ValueError("an image point has alpha^2 + beta^2 >= r_o^2: no direction on the observer's sky")
    if flags & KR_FLAG_EDGE:
        raise ValueError("the centre lies outside the captured region")
    if flags & KR_FLAG_DISK:
        raise ValueError("a disk crossing cannot be shaded: not equatorial, no circular orbit at its radius, or g <= 0")
    if flags & KR_FLAG_NOT_NULL:
        raise KrError(KR_ERR_NOT_NULL, "the ZAMO tetrad projection is not null to 1e-10, or no null p_t exists")
    if flags:
        raise KrError(KR_ERR_DEVICE, f"unknown device flags {flags:#x}")


class Handle:
    """One ``kr_handle`` (device, stream, scratch buffer); closed on garbage collection."""

    def __init__(self, device: int = 0) -> None:
        L = lib()
        h = ctypes.c_void_p()
        check(L.kr_create(device, ctypes.byref(h)), "kr_create")
        self._h = h
        self.device = device
        self._lock = threading.Lock()

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().kr_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass
