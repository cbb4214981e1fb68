"""gpu_tracking_amd -- gpu_tracking's per-frame particle detection on MI355X.

Hand-written HIP kernels for gfx950 behind a C ABI (include/gtrack.h,
libgtrack.so), with the reference's Python entry points on top:

    import gpu_tracking_amd as gt
    df = gt.batch(video, 9, minmass=500)

The package needs the built library; importing the entry points raises when
it is missing (no CPU fallback).
"""
from .api import (Batch, CentroidNet, Context, LoG, batch, characterize_points,  # noqa: F401
                  detections_to_numpy, link, make_log_params, make_params)
from ._native import GtError, column_names, default_params  # noqa: F401

__all__ = ["batch", "characterize_points", "LoG", "link", "Context", "Batch", "CentroidNet",
           "make_params",
           "make_log_params", "GtError",
           "column_names", "default_params", "detections_to_numpy"]
