"""ctypes binding of libgtrack.so (the C ABI declared in include/gtrack.h).

The library is built in-tree (``gpu_tracking_amd/libgtrack.so``) by
``__graft_entry__.build()`` / ``make -C gpu_tracking_amd/csrc``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgtrack.so")

GT_U16, GT_F32, GT_U8 = 0, 1, 2
GT_STYLE_TRACKPY, GT_STYLE_LOG = 0, 1

# gt_status codes (include/gtrack.h), one per variant of the reference's
# error enum (src/error.rs:6-94) plus two additions.
STATUS_NAMES = {
    0: "GT_OK", 1: "GpuAdapterError", 2: "GpuDeviceError", 3: "FileNotFound",
    4: "DimensionMismatch", 5: "EmptyIterator", 6: "NonSortedCharacterization",
    7: "FrameOutOfBounds", 8: "FrameOOB", 9: "ThreadError", 10: "NonStandardArrayLayout",
    11: "NoExtensionError", 12: "UnsupportedFileformat", 13: "ArrayDimensionsError",
    14: "CastError", 15: "ReadError", 16: "InvalidFileName", 17: "ChannelNotFound",
    18: "TooDenseToLink", 19: "Interrupted", 20: "PolledAfterTermination", 21: "TiffWrite",
    22: "WrongBuilder", 100: "InvalidArgument", 101: "CapacityExceeded",
}


class GtParams(C.Structure):
    """gt_params: TrackingParams + ParamStyle field for field (src/gpu_setup.rs:10-59)."""
    _fields_ = [
        ("style", C.c_int32),
        ("separation", C.c_uint32), ("diameter", C.c_uint32), ("maxsize", C.c_float),
        ("threshold", C.c_float), ("invert", C.c_int32), ("percentile", C.c_float),
        ("topn", C.c_uint32), ("preprocess", C.c_int32), ("filter_close", C.c_int32),
        ("min_radius", C.c_float), ("max_radius", C.c_float), ("n_radii", C.c_uint32),
        ("log_spacing", C.c_int32), ("overlap_threshold", C.c_float),
        ("minmass", C.c_float), ("max_iterations", C.c_uint32), ("characterize", C.c_int32),
        ("has_search_range", C.c_int32), ("search_range", C.c_float),
        ("has_memory", C.c_int32), ("memory", C.c_uint32),
        ("doughnut_correction", C.c_int32),
        ("has_bg_radius", C.c_int32), ("bg_radius", C.c_float),
        ("has_gap_radius", C.c_int32), ("gap_radius", C.c_float),
        ("has_snr", C.c_int32), ("snr", C.c_float),
        ("has_minmass_snr", C.c_int32), ("minmass_snr", C.c_float),
        ("truncate_preprocessed", C.c_int32),
        ("has_illumination_sigma", C.c_int32), ("illumination_sigma", C.c_float),
        ("has_adaptive_background", C.c_int32), ("adaptive_background", C.c_uint32),
        ("include_r_in_output", C.c_int32),
        ("shift_threshold", C.c_float),
        ("noise_size", C.c_float),
        ("has_smoothing_size", C.c_int32), ("smoothing_size", C.c_uint32),
        ("illumination_correction_per_frame", C.c_int32),
        ("threshold_rel", C.c_float),   # no reference counterpart (LoG; 0 = off)
    ]


class GtResultRow(C.Structure):
    """ResultRow (src/execute_gpu.rs:48-61)."""
    _fields_ = [(n, C.c_float) for n in
                ("x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc",
                 "count")]


class GtDetection(C.Structure):
    _fields_ = [("row", GtResultRow), ("seed", C.c_uint32), ("frame", C.c_uint32)]


class GtOutput(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("nrows", C.c_uint64), ("ncols", C.c_uint32),
                ("col_names", C.POINTER(C.c_char_p)), ("col_types", C.POINTER(C.c_char_p))]


# numpy view of gt_detection (48 bytes)
class GtFrameDesc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nrows", C.c_uint32), ("ncols", C.c_uint32),
                ("frame", C.c_uint32), ("pad", C.c_uint32)]


DETECTION_FIELDS = [
    ("x", "<f4"), ("y", "<f4"), ("mass", "<f4"), ("r", "<f4"), ("max_intensity", "<f4"),
    ("Rg", "<f4"), ("raw_mass", "<f4"), ("signal", "<f4"), ("ecc", "<f4"), ("count", "<f4"),
    ("seed", "<u4"), ("frame", "<u4"),
]

# Exported symbols, in header order (checked by the CPU test suite).
EXPORTS = [
    "gt_params_default", "gt_abi_version", "gt_last_error", "gt_context_create",
    "gt_context_destroy", "gt_detect_device", "gt_submit_frame", "gt_fetch_frame",
    "gt_context_set_inflight", "gt_detect_device_static", "gt_execute", "gt_output_free",
    "gt_column_names", "gt_filter_close", "gt_doughnut", "gt_set_interrupt", "gt_progress",
    "gt_set_timing", "gt_stage_times", "gt_prune_blobs", "gt_link", "gt_connect", "gt_batch_create",
    "gt_batch_destroy", "gt_batch_detect", "gt_batch_tune", "gt_batch_get_schedule",
    "gt_batch_get_schedule_info", "gt_batch_set_graph", "gt_frame_std", "gt_cnn_create", "gt_cnn_destroy", "gt_cnn_set_params",
    "gt_cnn_get_params", "gt_cnn_adam_steps", "gt_cnn_forward", "gt_cnn_gradient", "gt_cnn_step",
    "gt_cnn_train", "gt_cnn_make_samples", "gt_context_set_cnn", "gt_context_cnn_stats",
    "gt_file_open", "gt_file_close", "gt_file_info", "gt_file_read", "gt_file_channels",
    "gt_file_mean", "gt_execute_file", "gt_write_tiff", "gt_batch_set_lanes", "gt_batch_get_lanes",
    "gt_context_set_lanes", "gt_context_set_overlap",
]

CNN_NPARAM = 14147
CNN_PATCH_PITCH = 128

_lib = None


class GtError(BaseException):
    """Base of every error the library raises; `status` is the gt_status code.

    Each status is raised as a subclass that also derives from the built-in
    exception the reference's bindings raise for that error variant
    (gpu_tracking_python/src/python_bindings.rs:159-192), so code written
    against the reference (``except ValueError``, ``except
    FileNotFoundError``, Ctrl-C as KeyboardInterrupt) keeps working.  GtError
    itself derives from BaseException only so that the interrupt class stays
    outside ``except Exception``, as KeyboardInterrupt does.
    """

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class GtInterrupted(GtError, KeyboardInterrupt):
    """Error::Interrupted -> KeyboardInterrupt."""


class GtConnectionError(GtError, ConnectionError):
    """Error::GpuAdapterError / GpuDeviceError -> ConnectionError."""


class GtBaseError(GtError):
    """Error::ThreadError / PolledAfterTermination / TiffWrite -> BaseException."""


class GtFileNotFoundError(GtError, FileNotFoundError):
    """Error::FileNotFound / InvalidFileName -> FileNotFoundError."""


class GtValueError(GtError, ValueError):
    """Every other variant -> ValueError; also the two codes without a
    reference counterpart (rejected argument, capacity exceeded)."""


_STATUS_CLASS = {
    19: GtInterrupted,
    1: GtConnectionError, 2: GtConnectionError,
    9: GtBaseError, 20: GtBaseError, 21: GtBaseError,
    3: GtFileNotFoundError, 16: GtFileNotFoundError,
}


def error(status: int, message: str) -> GtError:
    """The exception object for a gt_status code (see GtError)."""
    return _STATUS_CLASS.get(status, GtValueError)(status, message)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"gpu_tracking_amd: the HIP library {LIB_PATH} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C gpu_tracking_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.gt_params_default.argtypes = [P(GtParams)]
    L.gt_params_default.restype = None
    L.gt_abi_version.restype = C.c_int32
    L.gt_last_error.restype = C.c_char_p
    L.gt_context_create.argtypes = [P(GtParams), C.c_uint32, C.c_uint32, C.c_int32,
                                    P(C.c_void_p)]
    L.gt_context_create.restype = C.c_int
    L.gt_context_destroy.argtypes = [C.c_void_p]
    L.gt_context_destroy.restype = None
    L.gt_detect_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_void_p,
                                   C.c_uint32, C.c_void_p, C.c_void_p]
    L.gt_detect_device.restype = C.c_int
    L.gt_submit_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_uint64,
                                  C.c_void_p, C.c_uint32]
    L.gt_submit_frame.restype = C.c_int
    L.gt_fetch_frame.argtypes = [C.c_void_p, P(C.c_uint64), C.c_void_p, C.c_uint32,
                                 P(C.c_uint32)]
    L.gt_fetch_frame.restype = C.c_int
    L.gt_context_set_inflight.argtypes = [C.c_void_p, C.c_uint32]
    L.gt_context_set_inflight.restype = C.c_int
    L.gt_detect_device_static.argtypes = [P(GtParams), C.c_uint32, C.c_uint32, C.c_void_p,
                                          C.c_int32, C.c_uint32, C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_void_p]
    L.gt_detect_device_static.restype = C.c_int
    L.gt_execute.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p,
                             C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32,
                             C.c_int32, C.c_void_p, C.c_uint64, P(GtOutput)]
    L.gt_execute.restype = C.c_int
    L.gt_output_free.argtypes = [P(GtOutput)]
    L.gt_output_free.restype = None
    L.gt_column_names.argtypes = [P(GtParams), P(C.c_char_p), P(C.c_char_p)]
    L.gt_column_names.restype = C.c_uint32
    L.gt_filter_close.argtypes = [C.c_void_p, C.c_uint64, C.c_float, C.c_void_p]
    L.gt_filter_close.restype = C.c_int
    L.gt_doughnut.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64,
                              C.c_float, C.c_float, C.c_float, C.c_void_p]
    L.gt_doughnut.restype = C.c_int
    L.gt_prune_blobs.argtypes = [C.c_void_p, C.c_uint64, C.c_float, C.c_void_p]
    L.gt_prune_blobs.restype = C.c_int
    L.gt_link.argtypes = [C.c_void_p, C.c_uint64, C.c_float, C.c_uint32, C.c_void_p]
    L.gt_link.restype = C.c_int
    L.gt_connect.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_float,
                             C.c_void_p, C.c_void_p]
    L.gt_connect.restype = C.c_int
    L.gt_batch_create.argtypes = [P(GtParams), C.c_uint64, C.c_uint64, C.c_uint32, P(C.c_void_p)]
    L.gt_batch_create.restype = C.c_int
    L.gt_batch_destroy.argtypes = [C.c_void_p]
    L.gt_batch_destroy.restype = None
    L.gt_batch_detect.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p,
                                  C.c_uint32, C.c_void_p, C.c_void_p]
    L.gt_batch_detect.restype = C.c_int
    L.gt_batch_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p,
                                C.c_uint32, C.c_void_p, P(C.c_float)]
    L.gt_batch_tune.restype = C.c_int
    L.gt_batch_get_schedule.argtypes = [C.c_void_p, C.c_void_p, P(C.c_uint64)]
    L.gt_batch_get_schedule.restype = C.c_int
    L.gt_batch_set_graph.argtypes = [C.c_void_p, C.c_int32]
    L.gt_batch_set_graph.restype = C.c_int
    L.gt_set_interrupt.argtypes = [C.c_void_p, C.c_int32]
    L.gt_set_interrupt.restype = C.c_int
    L.gt_progress.argtypes = [C.c_void_p, P(C.c_uint64), P(C.c_uint64)]
    L.gt_progress.restype = C.c_int
    L.gt_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.gt_set_timing.restype = C.c_int
    L.gt_stage_times.argtypes = [C.c_void_p, P(C.c_double), C.c_int32, P(C.c_uint32),
                                 P(C.c_uint32)]
    L.gt_stage_times.restype = C.c_int32
    vp, u64, u32, f32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_float
    sigs = {
        "gt_cnn_create": [u64, P(vp)],
        "gt_cnn_set_params": [vp, vp],
        "gt_cnn_get_params": [vp, vp],
        "gt_cnn_adam_steps": [vp, P(u64)],
        "gt_cnn_forward": [vp, vp, u64, vp, vp],
        "gt_cnn_gradient": [vp, vp, vp, vp, u32, vp, P(f32)],
        "gt_cnn_step": [vp, vp, vp, vp, u32, f32, P(f32)],
        "gt_cnn_train": [vp, vp, vp, u64, u32, u32, f32, u64, vp],
        "gt_cnn_make_samples": [vp, vp, C.c_int32, u32, vp, vp, u64, P(u64), P(u64), vp],
        "gt_context_set_cnn": [vp, vp, f32, f32, C.c_int32],
        "gt_frame_std": [vp, vp, C.c_int32, u32, vp, vp],
        "gt_context_cnn_stats": [vp, P(u64), P(u64), C.c_int32],
        "gt_file_open": [C.c_char_p, C.c_int32, u32, P(vp)],
        "gt_file_info": [vp, P(u32), P(u32), P(u64), P(C.c_int32)],
        "gt_file_read": [vp, u64, u64, vp],
        "gt_file_mean": [vp, vp],
        "gt_execute_file": [vp, vp, vp, u64, vp, u64, u32, C.c_int32, C.c_int32, vp, u64,
                            P(GtOutput)],
        "gt_write_tiff": [C.c_char_p, vp, u64, u32, u32],
        "gt_batch_set_lanes": [vp, u32],
        "gt_context_set_lanes": [vp, C.c_int32],
        "gt_context_set_overlap": [vp, C.c_int32],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.gt_cnn_destroy.argtypes = [vp]
    L.gt_cnn_destroy.restype = None
    L.gt_file_close.argtypes = [vp]
    L.gt_file_close.restype = None
    L.gt_file_channels.argtypes = [vp, vp, u32]
    L.gt_file_channels.restype = u32
    L.gt_batch_get_lanes.argtypes = [vp]
    L.gt_batch_get_lanes.restype = u32
    L.gt_batch_get_schedule_info.argtypes = [vp, vp]
    L.gt_batch_get_schedule_info.restype = C.c_int
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        msg = lib().gt_last_error().decode(errors="replace")
        raise error(status, msg)


def default_params() -> GtParams:
    p = GtParams()
    lib().gt_params_default(C.byref(p))
    return p


def column_names(p: GtParams):
    names = (C.c_char_p * 16)()
    types = (C.c_char_p * 16)()
    n = lib().gt_column_names(C.byref(p), names, types)
    return [names[i].decode() for i in range(n)], [types[i].decode() for i in range(n)]
