"""gpu_tracking_amd -- gpu_tracking's per-frame particle detection on MI355X.

Hand-written HIP kernels for gfx950 behind a C ABI (include/gtrack.h,
libgtrack.so), with the reference's Python entry points on top:

    import gpu_tracking_amd as gt
    df = gt.batch(video, 9, minmass=500)

The package needs the built library; importing the entry points raises when
it is missing (no CPU fallback).
"""
from .api import (Batch, CentroidNet, Context, LoG, VideoFile, batch,  # noqa: F401
                  characterize_points, connect, detect_device_static, detections_to_numpy, link, load, make_log_params,
                  make_params, mean_from_file, parse_ets, parse_ets_with_keys, to_tiff,
                  unique_particle_ids)
from ._native import (GtBaseError, GtConnectionError, GtError, GtFileNotFoundError,  # noqa: F401
                      GtInterrupted, GtValueError, column_names, default_params)

__all__ = ["batch", "characterize_points", "LoG", "link", "connect", "Context", "Batch", "CentroidNet",
           "VideoFile", "load", "mean_from_file", "parse_ets", "parse_ets_with_keys", "to_tiff",
           "make_params", "unique_particle_ids",
           "make_log_params", "GtError", "GtValueError", "GtFileNotFoundError",
           "GtConnectionError", "GtInterrupted", "GtBaseError",
           "column_names", "default_params", "detections_to_numpy", "detect_device_static"]
