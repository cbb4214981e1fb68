"""CPU oracle for the detection pipeline -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``oracle/liboracle.so`` (built from ``gt_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this module, and only to check
the HIP path against it.  The product package never imports it.

Functions mirror the reference WGSL/Rust they restate (see gt_oracle.h for
file:line citations).  Arrays are numpy; frames are (nrows, ncols) float32.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Params(C.Structure):
    _fields_ = [
        ("nrows", C.c_int32), ("ncols", C.c_int32), ("smoothing", C.c_int32),
        ("sigma2", C.c_float), ("circle", C.c_int32), ("dilation", C.c_int32),
        ("max_iterations", C.c_uint32), ("threshold", C.c_float), ("minmass", C.c_float),
        ("margin", C.c_int32), ("snr", C.c_float), ("minmass_snr", C.c_float),
        ("rough_snr_factor", C.c_float), ("image_std", C.c_float), ("truncate", C.c_int32),
        ("threshold_rel", C.c_float),
    ]


ROW_DTYPE = np.dtype([
    ("x", "<f4"), ("y", "<f4"), ("mass", "<f4"), ("r", "<f4"), ("max_intensity", "<f4"),
    ("Rg", "<f4"), ("raw_mass", "<f4"), ("signal", "<f4"), ("ecc", "<f4"), ("count", "<f4"),
    ("seed", "<u4"), ("pad", "<u4"),
])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.or_derive_trackpy.argtypes = [C.c_uint32, C.c_uint32, C.c_int32, C.c_float, C.c_float,
                                        C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_int,
                                        C.c_int, C.c_int32, C.c_int32, P(Params)]
        L.or_gauss_weights.argtypes = [C.c_int32, C.c_float, C.c_void_p, C.c_void_p]
        L.or_bandpass.argtypes = [C.c_void_p, P(Params), C.c_void_p, C.c_void_p]
        L.or_local_max.argtypes = [C.c_void_p, P(Params), C.c_void_p, C.c_size_t, C.c_void_p]
        L.or_local_max.restype = C.c_size_t
        L.or_walk.argtypes = [C.c_void_p, P(Params), C.c_void_p, C.c_size_t, C.c_void_p]
        L.or_walk.restype = C.c_size_t
        L.or_characterize.argtypes = [C.c_void_p, C.c_void_p, P(Params), C.c_void_p, C.c_size_t,
                                      C.c_int, C.c_int]
        L.or_filter_close.argtypes = [C.c_void_p, C.c_size_t, C.c_float, C.c_void_p]
        L.or_doughnut.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t,
                                  C.c_float, C.c_float, C.c_float, C.c_void_p]
        L.or_sort_rows.argtypes = [C.c_void_p, C.c_size_t]
        L.or_detect_frame.argtypes = [C.c_void_p, P(Params), C.c_int, C.c_void_p, C.c_size_t,
                                      C.c_void_p]
        L.or_detect_frame.restype = C.c_size_t
        L.or_detect_batch.argtypes = [C.c_void_p, C.c_int, C.c_size_t, P(Params), C.c_int, C.c_int,
                                      C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
        L.or_detect_batch.restype = C.c_int
        L.or_bandpass_f64.argtypes = [C.c_void_p, P(Params), C.c_void_p, C.c_void_p]
        L.or_centroid_f64.argtypes = [C.c_void_p, P(Params), C.c_int32, C.c_int32, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        L.or_walk_r.argtypes = [C.c_void_p, P(Params), C.c_void_p, C.c_void_p, C.c_size_t,
                                C.c_void_p]
        L.or_walk_r.restype = C.c_size_t
        L.or_log_radii.argtypes = [C.c_float, C.c_float, C.c_uint32, C.c_int, C.c_void_p]
        L.or_log_taps.argtypes = [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_log_taps.restype = C.c_int32
        L.or_log_plane.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
        L.or_log_max.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                 C.c_int32, C.c_float, C.c_void_p, C.c_size_t]
        L.or_log_max.restype = C.c_size_t
        L.or_blob_overlap.argtypes = [C.c_float, C.c_float, C.c_float]
        L.or_blob_overlap.restype = C.c_float
        L.or_prune_blobs.argtypes = [C.c_void_p, C.c_size_t, C.c_float]
        L.or_prune_blobs.restype = C.c_size_t
        L.or_log_detect_frame.argtypes = [C.c_void_p, P(Params), C.c_void_p, C.c_uint32, C.c_int,
                                          C.c_float, C.c_void_p, C.c_size_t, C.c_void_p]
        L.or_log_detect_frame.restype = C.c_size_t
        L.or_frame_std.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
        L.or_frame_std_div.argtypes = [C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint32,
                                       C.c_void_p, C.c_void_p]
        L.or_detect_frame_adaptive.argtypes = [C.c_void_p, P(Params), C.c_uint32, C.c_int,
                                               C.c_void_p, C.c_size_t]
        L.or_detect_frame_adaptive.restype = C.c_size_t
        L.or_log_detect_frame_adaptive.argtypes = [C.c_void_p, P(Params), C.c_void_p,
                                                   C.c_uint32, C.c_uint32, C.c_int, C.c_float,
                                                   C.c_void_p, C.c_size_t]
        L.or_log_detect_frame_adaptive.restype = C.c_size_t
        L.or_illum_taps.argtypes = [C.c_float, C.c_void_p]
        L.or_illum_taps.restype = C.c_int32
        L.or_illum_smooth.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p,
                                      C.c_void_p]
        L.or_frames_mean.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
        L.or_illum_divide.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def derive_trackpy(nrows: int, ncols: int, diameter: int = 9, separation: int | None = None,
                   smoothing_size: int | None = None, noise_size: float = 1.0,
                   minmass: float = 0.0, max_iterations: int = 10, shift_threshold: float = 0.6,
                   snr: float = 0.0, minmass_snr: float = 0.0, adaptive_background: bool = False,
                   truncate_preprocessed: bool = True, image_std: float = 0.0) -> Params:
    """gpu_setup.rs:528-580.  separation defaults to diameter + 1 as in the
    Python entry point (gpu_tracking_macros/src/lib.rs:508)."""
    p = Params()
    sep = diameter + 1 if separation is None else separation
    lib().or_derive_trackpy(diameter, sep, smoothing_size or 0, noise_size, minmass,
                            max_iterations, shift_threshold, snr, minmass_snr,
                            int(adaptive_background), int(truncate_preprocessed), nrows, ncols,
                            C.byref(p))
    p.image_std = image_std
    return p


def gauss_weights(s: int, sigma2: float):
    r = s // 2
    w = np.zeros(2 * r + 1, np.float32)
    n = np.zeros(1, np.float32)
    lib().or_gauss_weights(s, sigma2, _ptr(w), _ptr(n))
    return w, n[0]


def _f32(frame) -> np.ndarray:
    return np.ascontiguousarray(frame, dtype=np.float32)


def bandpass(frame, p: Params) -> np.ndarray:
    f = _f32(frame)
    out = np.empty_like(f)
    scratch = np.empty(2 * f.size, np.float32)
    lib().or_bandpass(_ptr(f), C.byref(p), _ptr(out), _ptr(scratch))
    return out


def local_max(processed, p: Params) -> np.ndarray:
    P_ = _f32(processed)
    cand = np.empty(P_.size, np.uint32)
    M = np.empty(P_.size, np.float32)
    n = lib().or_local_max(_ptr(P_), C.byref(p), _ptr(cand), cand.size, _ptr(M))
    return cand[:n].copy()


def walk(processed, p: Params, cand) -> np.ndarray:
    P_ = _f32(processed)
    cand = np.ascontiguousarray(cand, np.uint32)
    rows = np.zeros(max(len(cand), 1), ROW_DTYPE)
    n = lib().or_walk(_ptr(P_), C.byref(p), _ptr(cand), len(cand), _ptr(rows))
    return rows[:n].copy()


def characterize(processed, raw, p: Params, rows, use_f64: bool = True,
                 recompute_mass: bool = False) -> np.ndarray:
    P_ = _f32(processed)
    R = _f32(raw)
    rows = np.ascontiguousarray(rows, ROW_DTYPE).copy()
    lib().or_characterize(_ptr(P_), _ptr(R), C.byref(p), _ptr(rows), len(rows), int(use_f64),
                          int(recompute_mass))
    return rows


def filter_close(rows, separation: float) -> np.ndarray:
    rows = np.ascontiguousarray(rows, ROW_DTYPE)
    keep = np.zeros(max(len(rows), 1), np.uint8)
    lib().or_filter_close(_ptr(rows), len(rows), separation, _ptr(keep))
    return keep[:len(rows)].astype(bool)


def doughnut(raw, rows, radius: float, bg_radius: float, gap_radius: float) -> np.ndarray:
    R = _f32(raw)
    rows = np.ascontiguousarray(rows, ROW_DTYPE)
    out = np.zeros((max(len(rows), 1), 3), np.float32)
    lib().or_doughnut(_ptr(R), R.shape[0], R.shape[1], _ptr(rows), len(rows), radius, bg_radius,
                      gap_radius, _ptr(out))
    return out[:len(rows)]


def detect_frame(frame, p: Params, characterize: bool = True):
    """bandpass -> local max -> walk -> characterize (f64 moments); rows are in
    candidate (raster) order.  Returns (rows, n_candidates)."""
    f = _f32(frame)
    cap = f.size
    rows = np.zeros(cap, ROW_DTYPE)
    nc = C.c_size_t(0)
    n = lib().or_detect_frame(_ptr(f), C.byref(p), int(characterize), _ptr(rows), cap,
                              C.byref(nc))
    return rows[:n].copy(), nc.value


def detect_batch(frames: np.ndarray, p: Params, characterize: bool = True,
                 nthreads: int = 1, cap_per_frame: int | None = None):
    """Frame-parallel oracle over a (T, H, W) uint16 or float32 stack.
    Returns (list of per-frame row arrays, candidate counts)."""
    frames = np.ascontiguousarray(frames)
    if frames.dtype == np.uint16:
        dt = 0
    else:
        frames = frames.astype(np.float32, copy=False)
        dt = 1
    T = frames.shape[0]
    cap = cap_per_frame or frames.shape[1] * frames.shape[2] // 4
    rows = np.zeros(T * cap, ROW_DTYPE)
    counts = np.zeros(T, np.uint64)
    ncand = np.zeros(T, np.uint64)
    st = lib().or_detect_batch(_ptr(frames), dt, T, C.byref(p), int(characterize), nthreads,
                               _ptr(rows), cap, _ptr(counts), _ptr(ncand))
    if st != 0:
        raise RuntimeError("oracle: detection capacity exceeded")
    return [rows[t * cap: t * cap + int(counts[t])].copy() for t in range(T)], ncand


def bandpass_f64(frame, p: Params) -> np.ndarray:
    f = _f32(frame)
    out = np.empty(f.shape, np.float64)
    scratch = np.empty(2 * f.size, np.float64)
    lib().or_bandpass_f64(_ptr(f), C.byref(p), _ptr(out), _ptr(scratch))
    return out


def centroid_f64(processed64, p: Params, pu: int, pv: int):
    P_ = np.ascontiguousarray(processed64, np.float64)
    dx, dy, m = C.c_double(), C.c_double(), C.c_double()
    lib().or_centroid_f64(_ptr(P_), C.byref(p), pu, pv, C.byref(dx), C.byref(dy), C.byref(m))
    return dx.value, dy.value, m.value


# ---------------------------------------------------------------------------
# LoG style


def derive_log(nrows: int, ncols: int, max_radius: float, smoothing_size: int | None = None,
               noise_size: float = 1.0, minmass: float = 0.0, max_iterations: int = 10,
               shift_threshold: float = 0.6, truncate_preprocessed: bool = True,
               threshold_rel: float = 0.0) -> Params:
    """gpu_setup.rs:549-557 + :562-579 (Log branch): smoothing side
    2 * u32(max_radius + 0.5) + 1 unless given; no circle, dilation or margin."""
    p = Params()
    if smoothing_size is None:
        rad = int(np.float32(max_radius) + np.float32(0.5)) if max_radius > -0.5 else 0
        smoothing_size = 2 * rad + 1
    p.nrows, p.ncols = nrows, ncols
    p.smoothing = smoothing_size
    p.sigma2 = noise_size
    p.circle = 0
    p.dilation = 0
    p.max_iterations = max_iterations
    p.threshold = shift_threshold
    p.minmass = minmass
    p.margin = 0
    p.snr = 0.0
    p.minmass_snr = 0.0
    p.rough_snr_factor = 1.0
    p.image_std = 0.0
    p.truncate = int(truncate_preprocessed)
    p.threshold_rel = threshold_rel
    return p


def log_radii(min_r: float, max_r: float, n: int, log_spacing: bool) -> np.ndarray:
    out = np.zeros(max(n, 1), np.float32)
    lib().or_log_radii(min_r, max_r, n, int(log_spacing), _ptr(out))
    return out[:n]


def log_taps(sigma: float):
    """(R, g, d, norm1, norm3) of one LoG scale."""
    R = lib().or_log_taps(sigma, None, None, None, None)
    g = np.zeros(2 * R + 1, np.float32)
    d = np.zeros(2 * R + 1, np.float32)
    n1, n3 = C.c_float(), C.c_float()
    lib().or_log_taps(sigma, _ptr(g), _ptr(d), C.byref(n1), C.byref(n3))
    return R, g, d, n1.value, n3.value


def log_plane(raw, sigma: float) -> np.ndarray:
    f = _f32(raw)
    L = np.empty_like(f)
    t1 = np.empty_like(f)
    t2 = np.empty_like(f)
    lib().or_log_plane(_ptr(f), f.shape[0], f.shape[1], sigma, _ptr(L), _ptr(t1), _ptr(t2))
    return L


def log_max(bottom, middle, top, processed, snr_threshold: float = 0.0) -> np.ndarray:
    M = _f32(middle)
    B = None if bottom is None else _f32(bottom)
    T = None if top is None else _f32(top)
    P_ = _f32(processed)
    cand = np.empty(M.size, np.uint32)
    n = lib().or_log_max(None if B is None else _ptr(B), _ptr(M), None if T is None else _ptr(T),
                         _ptr(P_), M.shape[0], M.shape[1], snr_threshold, _ptr(cand), cand.size)
    return cand[:n].copy()


def walk_r(processed, p: Params, cand, radii) -> np.ndarray:
    P_ = _f32(processed)
    cand = np.ascontiguousarray(cand, np.uint32)
    rr = np.ascontiguousarray(radii, np.float32)
    rows = np.zeros(max(len(cand), 1), ROW_DTYPE)
    n = lib().or_walk_r(_ptr(P_), C.byref(p), _ptr(cand), _ptr(rr), len(cand), _ptr(rows))
    return rows[:n].copy()


def blob_overlap(r1: float, r2: float, distance: float) -> float:
    return lib().or_blob_overlap(r1, r2, distance)


def prune_blobs(rows, overlap_threshold: float) -> np.ndarray:
    rows = np.ascontiguousarray(rows, ROW_DTYPE).copy()
    n = lib().or_prune_blobs(_ptr(rows), len(rows), overlap_threshold)
    return rows[:n].copy()


def log_detect_frame(frame, p: Params, radii, characterize: bool = True,
                     overlap_threshold: float = 0.0):
    """LoG frame: rows in the pruned (r desc, mass desc) order.  Returns
    (rows, n_candidates); seed = scale * H * W + raster index."""
    f = _f32(frame)
    rr = np.ascontiguousarray(radii, np.float32)
    cap = f.size * max(len(rr), 1)
    rows = np.zeros(cap, ROW_DTYPE)
    nc = C.c_size_t(0)
    n = lib().or_log_detect_frame(_ptr(f), C.byref(p), _ptr(rr), len(rr), int(characterize),
                                  overlap_threshold, _ptr(rows), cap, C.byref(nc))
    return rows[:n].copy(), nc.value


def log_detect_frame_adaptive(frame, p: Params, radii, n_iter: int, characterize: bool = True,
                              overlap_threshold: float = 0.0):
    """LoG frame with adaptive_background = n_iter (p from derive_log with
    snr / minmass_snr / minmass set and rough_snr_factor 0.5): accepted rows,
    characterized and pruned."""
    f = _f32(frame)
    rr = np.ascontiguousarray(radii, np.float32)
    cap = f.size * max(len(rr), 1)
    rows = np.zeros(cap, ROW_DTYPE)
    n = lib().or_log_detect_frame_adaptive(_ptr(f), C.byref(p), _ptr(rr), len(rr), int(n_iter),
                                           int(characterize), overlap_threshold, _ptr(rows), cap)
    return rows[:n].copy()


def frame_std(frame):
    """(mean, std) noise estimate of a raw frame (f32 values; NaN-marker pixels
    with bits 0xFFFFFFFF are excluded from the sums but not the divisors)."""
    f = np.ascontiguousarray(frame, dtype=np.float32).reshape(-1)
    if f.size <= 24:
        raise ValueError("the noise estimate needs more than 24 pixels")
    m = np.zeros(1, np.float32)
    s = np.zeros(1, np.float32)
    lib().or_frame_std(_ptr(f), f.size, _ptr(m), _ptr(s))
    return float(m[0]), float(s[0])


def frame_std_div(frame, div_mean, div_var):
    """frame_std with explicit divisors (the adaptive filter's re-estimates)."""
    f = np.ascontiguousarray(frame, dtype=np.float32).reshape(-1)
    m = np.zeros(1, np.float32)
    s = np.zeros(1, np.float32)
    lib().or_frame_std_div(_ptr(f), f.size, int(div_mean), int(div_var), _ptr(m), _ptr(s))
    return float(m[0]), float(s[0])


def detect_frame_adaptive(frame, p: Params, n_iter: int, characterize: bool = True):
    """Trackpy detection with adaptive_background = n_iter (p from
    derive_trackpy(..., adaptive_background=True)); accepted rows."""
    f = _f32(frame)
    rows = np.zeros(max(f.size, 1), ROW_DTYPE)
    n = lib().or_detect_frame_adaptive(_ptr(f), C.byref(p), int(n_iter), int(characterize),
                                       _ptr(rows), rows.size)
    return rows[:n].copy()


def illum_taps(sigma: float) -> np.ndarray:
    """The illumination smoothing's taps (w[i + rint], i = -rint..rint)."""
    rint = int(np.float32(4.0) * np.float32(sigma) + np.float32(0.5))
    w = np.zeros(2 * rint + 1, np.float32)
    r = lib().or_illum_taps(C.c_float(sigma), _ptr(w))
    assert r == rint
    return w


def illum_smooth(frame, sigma: float) -> np.ndarray:
    """smooth_mean_frame's two passes over one f32 frame."""
    f = np.ascontiguousarray(frame, dtype=np.float32)
    H, W = f.shape
    out = np.zeros_like(f)
    tmp = np.zeros_like(f)
    lib().or_illum_smooth(_ptr(f), H, W, C.c_float(sigma), _ptr(out), _ptr(tmp))
    return out


def frames_mean(frames) -> np.ndarray:
    """mean_from_iter: the per-pixel sequential f32 mean over frames."""
    f = np.ascontiguousarray(frames, dtype=np.float32)
    out = np.zeros(f.shape[1:], np.float32)
    lib().or_frames_mean(_ptr(f), f.shape[0], out.size, _ptr(out))
    return out


def illum_correct(frames, sigma: float, per_frame: bool = False) -> np.ndarray:
    """The frames the detection sees with illumination correction: each raw
    frame (as f32) divided by the smoothed mean frame of the whole video, or
    by its own smoothed copy when per_frame."""
    f = np.ascontiguousarray(frames, dtype=np.float32)
    out = np.empty_like(f)
    prof = None if per_frame else illum_smooth(frames_mean(f), sigma)
    for t in range(f.shape[0]):
        pr = illum_smooth(f[t], sigma) if per_frame else prof
        lib().or_illum_divide(_ptr(f[t]), _ptr(pr), f[t].size, _ptr(out[t]))
    return out
