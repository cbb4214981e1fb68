"""Python entry points with the reference's names, arguments and defaults.

``batch``, ``characterize_points`` and ``LoG`` mirror gpu_tracking's
``python/gpu_tracking/generated.py`` (batch :2-57, characterize_points
:59-162, LoG :164-227) and the defaults its binding macro applies
(gpu_tracking_macros/src/lib.rs:187-683).  They return a pandas DataFrame
with the reference's columns (column_names, src/execute_gpu.rs:482-519) and
dtypes.  Frames are given as a numpy array (T, Y, X) or (Y, X); uint16,
uint8 and float32 frames go to the GPU without a host-side cast.

Fixes of reference wrapper bugs (documented divergences):
  - characterize_points: ``has_frames = None,`` made a tuple
    (generated.py:127) and the function called a missing
    ``characterize_rust`` (generated.py:154);
  - LoG passed undefined names ``min_r, max_r`` (generated.py:222-225).
"""
from __future__ import annotations

import ctypes as C
import threading
import os
import math
from typing import Iterable, Optional

import numpy as np

from . import _native as N

# Python-level defaults of the reference binding (macros lib.rs:187-461, 487-569).
_TRACKPY_DEFAULTS = dict(
    noise_size=1.0, smoothing_size=None, minmass=0.0, max_iterations=10, characterize=True,
    search_range=None, memory=None, doughnut_correction=True, bg_radius=None, gap_radius=None,
    snr=None, minmass_snr=None, truncate_preprocessed=True, correct_illumination=None,
    illumination_sigma=None, illumination_correction_per_frame=False, adaptive_background=None,
    shift_threshold=0.6, linker_reset_points=None, frames=None, tqdm=True, channel=None,
    maxsize=math.inf, separation=None, threshold=1.0 / 255.0, invert=False, percentile=64.0,
    topn=0xFFFFFFFF, preprocess=True, filter_close=True,
)


def _dtype_code(arr: np.ndarray) -> int:
    if arr.dtype == np.uint16:
        return N.GT_U16
    if arr.dtype == np.uint8:
        return N.GT_U8
    if arr.dtype == np.float32:
        return N.GT_F32
    raise TypeError(f"unsupported frame dtype {arr.dtype}")


_NP_OF_CODE = {N.GT_U8: np.uint8, N.GT_U16: np.uint16, N.GT_F32: np.float32}


def _is_path(x) -> bool:
    return isinstance(x, (str, bytes)) or hasattr(x, "__fspath__")


class VideoFile:
    """A video file read by the library (path_to_iter, src/execute_gpu.rs:
    1186-1270): tif / tiff (multi-page), ets, or vsi (whose frames are in
    _<stem>_/stack1/frame_t_0.ets beside it); `channel` picks the ETS
    channel (default 0).  Frames keep their sample type (u8, u16, f32; i8 and
    i16 read as f32).  shape = (frames, rows, cols)."""

    def __init__(self, path, channel=None):
        self._lib = N.lib()
        self._h = C.c_void_p()
        N.check(self._lib.gt_file_open(os.fsencode(path), int(channel is not None),
                                       int(channel or 0), C.byref(self._h)))
        nr, nc, nf, dt = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_int32()
        N.check(self._lib.gt_file_info(self._h, C.byref(nr), C.byref(nc), C.byref(nf),
                                       C.byref(dt)))
        self.shape = (int(nf.value), int(nr.value), int(nc.value))
        self.code = int(dt.value)
        self.dtype = np.dtype(_NP_OF_CODE[self.code])

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def __len__(self) -> int:
        return self.shape[0]

    def read(self, keys=None) -> np.ndarray:
        """Frames `keys` (default: all, in order) as a (n, rows, cols) array of
        the file's sample type."""
        idx = range(self.shape[0]) if keys is None else [int(k) for k in keys]
        out = np.empty((len(idx),) + self.shape[1:], self.dtype)
        # runs of consecutive frames go in one call
        i = 0
        idx = list(idx)
        while i < len(idx):
            j = i + 1
            while j < len(idx) and idx[j] == idx[j - 1] + 1:
                j += 1
            if idx[i] < 0:
                raise N.error(7, f"Requested a frame outside the length\n of the video. Video "
                                   f"length is {self.shape[0]}, requested frame {idx[i]}")
            N.check(self._lib.gt_file_read(self._h, idx[i], j - i, out[i:j].ctypes.data))
            i = j
        return out

    def mean(self) -> np.ndarray:
        """mean_from_iter (src/execute_gpu.rs:819-847): per-pixel f32 sum in
        frame order divided by the frame count."""
        out = np.empty(self.shape[1:], np.float32)
        N.check(self._lib.gt_file_mean(self._h, out.ctypes.data))
        return out

    def channels(self) -> list:
        """ETS channels present in the file (empty for TIFF)."""
        n = self._lib.gt_file_channels(self._h, None, 0)
        buf = (C.c_uint32 * max(int(n), 1))()
        self._lib.gt_file_channels(self._h, buf, n)
        return [int(buf[i]) for i in range(n)]

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.gt_file_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _as_video(file_or_array, channel=None):
    if isinstance(file_or_array, VideoFile):
        return file_or_array
    if _is_path(file_or_array):
        return VideoFile(file_or_array, channel)
    arr = np.asarray(file_or_array)
    if arr.ndim == 2:
        arr = arr[None]
    if arr.ndim != 3:
        raise ValueError(f"expected a (T, Y, X) array, got shape {arr.shape}")
    if arr.dtype not in (np.uint16, np.uint8, np.float32):
        # the reference's np.atleast_3d(arr).astype("float32") (python_bindings.rs:144)
        arr = arr.astype(np.float32)
    return np.ascontiguousarray(arr)


def make_params(diameter: int, **kw) -> N.GtParams:
    """TrackingParams for the Trackpy style with the Python binding's defaults."""
    unknown = set(kw) - set(_TRACKPY_DEFAULTS)
    if unknown:
        raise TypeError(f"unexpected keyword arguments: {sorted(unknown)}")
    a = dict(_TRACKPY_DEFAULTS)
    a.update({k: v for k, v in kw.items() if v is not None or k in ("search_range", "memory")})
    p = N.default_params()
    p.style = N.GT_STYLE_TRACKPY
    p.diameter = int(diameter)
    p.separation = int(a["separation"]) if a["separation"] is not None else int(diameter) + 1
    p.maxsize = float(a["maxsize"])
    p.threshold = float(a["threshold"])
    p.invert = int(bool(a["invert"]))
    p.percentile = float(a["percentile"])
    p.topn = int(a["topn"])
    p.preprocess = int(bool(a["preprocess"]))
    p.filter_close = int(bool(a["filter_close"]))
    _common(p, a)
    return p


def _common(p: N.GtParams, a: dict) -> None:
    p.noise_size = float(a["noise_size"])
    if a["smoothing_size"] is not None:
        p.has_smoothing_size, p.smoothing_size = 1, int(a["smoothing_size"])
    p.minmass = float(a["minmass"])
    p.max_iterations = int(a["max_iterations"])
    p.characterize = int(bool(a["characterize"]))
    if a["search_range"] is not None:
        p.has_search_range, p.search_range = 1, float(a["search_range"])
    if a["memory"] is not None:
        p.has_memory, p.memory = 1, int(a["memory"])
    p.doughnut_correction = int(bool(a["doughnut_correction"]))
    if a["bg_radius"] is not None:
        p.has_bg_radius, p.bg_radius = 1, float(a["bg_radius"])
        # gap_radius only takes effect together with bg_radius (macros lib.rs:312)
        p.has_gap_radius = 1
        p.gap_radius = float(a["gap_radius"]) if a["gap_radius"] is not None else 0.0
    if a["snr"] is not None:
        p.has_snr, p.snr = 1, float(a["snr"])
    if a["minmass_snr"] is not None:
        p.has_minmass_snr, p.minmass_snr = 1, float(a["minmass_snr"])
    p.truncate_preprocessed = int(bool(a["truncate_preprocessed"]))
    sigma = a["illumination_sigma"]
    if sigma is None and a.get("correct_illumination"):
        sigma = 30.0
    if sigma is not None:
        p.has_illumination_sigma, p.illumination_sigma = 1, float(sigma)
    p.illumination_correction_per_frame = int(bool(a["illumination_correction_per_frame"]))
    if a["adaptive_background"] is not None:
        p.has_adaptive_background, p.adaptive_background = 1, int(a["adaptive_background"])
    p.shift_threshold = float(a["shift_threshold"])


def _torch_code(frames) -> int:
    import torch
    if frames.dtype in (torch.int16, torch.uint16):
        return N.GT_U16
    if frames.dtype == torch.uint8:
        return N.GT_U8
    if frames.dtype == torch.float32:
        return N.GT_F32
    raise TypeError(f"unsupported dtype {frames.dtype}")


class Context:
    """One detection context (setup_state) for frames of a fixed size."""

    def __init__(self, params: N.GtParams, nrows: int, ncols: int,
                 characterize_points: bool = False):
        self._lib = N.lib()
        self.params = params
        self.shape = (int(nrows), int(ncols))
        h = C.c_void_p()
        N.check(self._lib.gt_context_create(C.byref(params), nrows, ncols,
                                            int(characterize_points), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.gt_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def execute(self, video: np.ndarray, keys: Optional[Iterable[int]] = None,
                points: Optional[np.ndarray] = None, points_has_frames: bool = False,
                points_has_r: bool = False, linker_reset_points=None):
        """gt_execute over a host video (or gt_execute_file over a VideoFile);
        returns (matrix, names, types)."""
        if not isinstance(video, VideoFile):
            video = np.ascontiguousarray(video)
        if tuple(video.shape[1:]) != self.shape:
            raise N.error(4, f"dimension mismatch. provided dimensions are {self.shape}, "
                               f"but the frames are {video.shape[1:]}")
        keys_arr = None if keys is None else np.ascontiguousarray(np.asarray(list(keys), np.uint64))
        pts = None if points is None else np.ascontiguousarray(points, np.float32)
        reset = None if linker_reset_points is None else np.ascontiguousarray(
            np.asarray(list(linker_reset_points), np.uint64))
        out = N.GtOutput()
        rest = (None if keys_arr is None else keys_arr.ctypes.data,
                0 if keys_arr is None else len(keys_arr),
                None if pts is None else pts.ctypes.data, 0 if pts is None else pts.shape[0],
                0 if pts is None else pts.shape[1], int(points_has_frames), int(points_has_r),
                None if reset is None else reset.ctypes.data, 0 if reset is None else len(reset),
                C.byref(out))
        if isinstance(video, VideoFile):
            st, interrupted, msg = self._interruptible(self._lib.gt_execute_file, self._h,
                                                  video.handle, *rest)
        else:
            st, interrupted, msg = self._interruptible(self._lib.gt_execute, self._h,
                                                  video.ctypes.data, _dtype_code(video),
                                                  video.shape[0], *rest)
        if interrupted and st == 0:
            # the call finished before it saw the flag: the interrupt still wins
            self._lib.gt_output_free(C.byref(out))
            raise N.error(19, "Keyboard Interrupt")
        if st != 0:
            raise N.error(st, msg)
        try:
            n, m = int(out.nrows), int(out.ncols)
            names = [out.col_names[i].decode() for i in range(m)]
            types = [out.col_types[i].decode() for i in range(m)]
            data = np.ctypeslib.as_array(out.data, shape=(n * m,)).copy().reshape(n, m) \
                if n * m else np.zeros((0, m), np.float32)
        finally:
            self._lib.gt_output_free(C.byref(out))
        return data, names, types

    def _interruptible(self, fn, *args):
        """Runs a whole-video call on a worker thread while this thread
        waits in 50 ms steps, so that Ctrl-C reaches Python during the call.
        A KeyboardInterrupt sets the context's interrupt flag; the call then
        ends with Error::Interrupted, raised as GtInterrupted (a
        KeyboardInterrupt).  The reference's bindings do the same with a
        worker thread polled every 50 ms (gpu_tracking_macros/src/lib.rs:75-129)."""
        box = {}

        def work():
            box["st"] = fn(*args)
            # the library keeps the message per thread: read it on this one
            box["msg"] = self._lib.gt_last_error().decode(errors="replace") if box["st"] else ""

        t = threading.Thread(target=work, name="gt_execute", daemon=True)
        t.start()
        interrupted = False
        try:
            while t.is_alive():
                t.join(0.05)
        except KeyboardInterrupt:
            interrupted = True
            self._lib.gt_set_interrupt(self._h, 1)
            t.join()
        return box["st"], interrupted, box["msg"]

    # one frame at a time (submit_work / get_work) ----------------------------
    def submit_frame(self, frame: np.ndarray, frame_idx: int, positions=None) -> None:
        """gt_submit_frame: stage one host frame (and, for characterize_points
        contexts, its positions: an (n, 2) [y, x] or (n, 3) [y, x, r] array)
        and queue its detection; returns without waiting."""
        frame = np.ascontiguousarray(frame)
        if tuple(frame.shape) != self.shape:
            raise N.error(4, f"dimension mismatch. provided dimensions are {self.shape}, "
                             f"but frame {frame_idx} has shape {frame.shape}")
        pos, npos, buf = None, 0, None
        if positions is not None:
            pa = np.asarray(positions, np.float32).reshape(-1, np.shape(positions)[-1])
            buf = np.zeros((len(pa), 10), np.float32)     # ResultRow: x = row, y = column
            buf[:, 0] = pa[:, 0]
            buf[:, 1] = pa[:, 1]
            if pa.shape[1] > 2:
                buf[:, 3] = pa[:, 2]
            pos, npos = buf.ctypes.data, len(pa)
        N.check(self._lib.gt_submit_frame(self._h, frame.ctypes.data, _dtype_code(frame),
                                          frame.size, int(frame_idx), pos, npos))

    def fetch_frame(self, cap: int = 4096):
        """gt_fetch_frame: (frame_idx, rows) of the oldest frame in flight;
        rows is a structured array of the ResultRow fields in ascending seed
        order (fetched again with a larger buffer when the frame has more
        than cap rows)."""
        idx, n = C.c_uint64(0), C.c_uint32(0)
        dt = np.dtype([(f, "<f4") for f, _ in N.GtResultRow._fields_])
        while True:
            out = np.zeros(max(1, cap), dt)
            st = self._lib.gt_fetch_frame(self._h, C.byref(idx), out.ctypes.data, cap,
                                          C.byref(n))
            if st == 101 and n.value > cap:
                cap = n.value
                continue
            N.check(st)
            return int(idx.value), out[:n.value].copy()

    def set_inflight(self, depth: int) -> None:
        """Frames submit_frame may have in flight (1..16, default 2)."""
        N.check(self._lib.gt_context_set_inflight(self._h, int(depth)))

    def stream(self, frames, positions=None, depth: int = 2):
        """Detection of an iterable of host frames one at a time, as
        execute_gpu's loop drives submit_work / get_work: up to `depth` frames
        are submitted before the oldest is fetched.  Yields (frame_idx, rows)
        in frame order.  positions: for characterize_points contexts, a
        callable frame_idx -> positions array."""
        self.set_inflight(depth)
        inflight = 0
        for k, fr in enumerate(frames):
            if inflight == depth:
                yield self.fetch_frame()
                inflight -= 1
            self.submit_frame(fr, k, None if positions is None else positions(k))
            inflight += 1
        while inflight:
            yield self.fetch_frame()
            inflight -= 1

    def set_lanes(self, lanes: int) -> None:
        """Streams (1 or 2) the chunks of a detect_device call alternate
        between (plain Trackpy-style detection); results do not depend on it."""
        N.check(self._lib.gt_context_set_lanes(self._h, int(lanes)))

    def set_overlap(self, enable: bool = True) -> None:
        """Runs each chunk's slow walks and characterize on a second stream,
        overlapped with the next chunk's bandpass and local maxima (plain
        Trackpy-style detection, one lane); results do not depend on it."""
        N.check(self._lib.gt_context_set_overlap(self._h, int(bool(enable))))

    # device-resident path -------------------------------------------------
    def detect_device(self, frames, out=None, counts=None, cap_per_frame: Optional[int] = None,
                      stream=None):
        """Runs the pipeline on frames already on the GPU (a torch tensor of
        shape (T, Y, X), dtype int16/uint16 (read as u16), uint8 or float32).
        Returns (detections tensor (T, cap, 12) of 32-bit words, counts (T,)
        int32); the call is asynchronous on `stream` (default: torch's current
        stream)."""
        import torch
        if not frames.is_cuda:
            raise ValueError("detect_device needs a device tensor")
        if frames.dim() == 2:
            frames = frames.unsqueeze(0)
        if tuple(frames.shape[1:]) != self.shape or not frames.is_contiguous():
            raise ValueError("frames must be contiguous with shape (T,) + context shape")
        code = _torch_code(frames)
        T = frames.shape[0]
        cap = cap_per_frame or max(1024, (self.shape[0] * self.shape[1]) // 16)
        if out is None:
            out = torch.empty((T, cap, 12), dtype=torch.int32, device=frames.device)
        if counts is None:
            counts = torch.empty((T,), dtype=torch.int32, device=frames.device)
        if stream is None:
            stream = torch.cuda.current_stream(frames.device).cuda_stream
        N.check(self._lib.gt_detect_device(self._h, frames.data_ptr(), code, T, out.data_ptr(),
                                           cap, counts.data_ptr(), stream))
        return out, counts

    def frame_std(self, frames, stream=None):
        """The noise estimate the snr / minmass_snr rules use: each raw
        frame's sample standard deviation (float32 device tensor (T,))."""
        import torch
        if frames.dim() == 2:
            frames = frames.unsqueeze(0)
        code = _torch_code(frames)
        if tuple(frames.shape[1:]) != self.shape or not frames.is_contiguous():
            raise ValueError("frames must be contiguous with shape (T,) + context shape")
        out = torch.empty((frames.shape[0],), dtype=torch.float32, device=frames.device)
        if stream is None:
            stream = torch.cuda.current_stream(frames.device).cuda_stream
        N.check(self._lib.gt_frame_std(self._h, frames.data_ptr(), code, frames.shape[0],
                                       out.data_ptr(), stream))
        return out

    def set_learned(self, net: Optional["CentroidNet"], conf_logit: float = 0.0,
                    tol: float = 0.05, scope: str = "moving") -> None:
        """Detect with the network's centroid proposals (None: the plain
        walk).  scope "moving": the network proposes for the candidates whose
        first walk step moves (a walk that stops at once is finished by that
        step); "all": for every candidate.  A proposal is kept when its logit
        >= conf_logit and one walk evaluation at its centre neither moves nor
        lands more than tol px away; every other candidate finishes the full
        walk."""
        scopes = {"moving": 0, "all": 1}
        if scope not in scopes:
            raise ValueError(f"scope must be one of {sorted(scopes)}")
        N.check(self._lib.gt_context_set_cnn(self._h, None if net is None else net.handle,
                                             float(conf_logit), float(tol), scopes[scope]))
        self._net = net

    def learned_stats(self, reset: bool = False):
        """(accepted proposals, candidates) since the last reset."""
        a, c = C.c_uint64(0), C.c_uint64(0)
        N.check(self._lib.gt_context_cnn_stats(self._h, C.byref(a), C.byref(c), int(reset)))
        return a.value, c.value

    def training_samples(self, frames, cap: Optional[int] = None, stream=None):
        """Training pairs from device frames (T, Y, X): every candidate's
        normalised patch (bfloat16 (n, 128), entries 121.. zero) and target
        (float32 (n, 4): end row - candidate row, end column - candidate
        column, 1 if the walk converges, candidate raster index as uint32
        bits) from the plain method's full walk.
        Returns (patches, targets, candidates seen)."""
        import torch
        if frames.dim() == 2:
            frames = frames.unsqueeze(0)
        code = _torch_code(frames)
        if tuple(frames.shape[1:]) != self.shape or not frames.is_contiguous():
            raise ValueError("frames must be contiguous with shape (T,) + context shape")
        T = frames.shape[0]
        if cap is None:
            cap = T * max(1024, (self.shape[0] * self.shape[1]) // 32)
        patches = torch.zeros((cap, N.CNN_PATCH_PITCH), dtype=torch.bfloat16, device=frames.device)
        targets = torch.zeros((cap, 4), dtype=torch.float32, device=frames.device)
        nw, nt = C.c_uint64(0), C.c_uint64(0)
        if stream is None:
            stream = torch.cuda.current_stream(frames.device).cuda_stream
        N.check(self._lib.gt_cnn_make_samples(self._h, frames.data_ptr(), code, T,
                                              patches.data_ptr(), targets.data_ptr(), cap,
                                              C.byref(nw), C.byref(nt), stream))
        n = nw.value
        return patches[:n], targets[:n], nt.value

    def set_timing(self, enable: bool) -> None:
        N.check(self._lib.gt_set_timing(self._h, int(enable)))

    def stage_times(self):
        """(per-stage ms of the last detect_device call, launches per stage,
        frames per launch); stages are bandpass, local max, walk."""
        buf = (C.c_double * 3)()
        nl, fpl = C.c_uint32(0), C.c_uint32(0)
        n = self._lib.gt_stage_times(self._h, buf, 3, C.byref(nl), C.byref(fpl))
        return [buf[i] for i in range(n)], nl.value, fpl.value


def detect_device_static(params: N.GtParams, frames, out=None, counts=None,
                         cap_per_frame: Optional[int] = None, stream=None):
    """gt_detect_device_static: Context.detect_device's plain path without a
    context, its scratch in the library's static device array (no device
    allocation inside the call; `out` and `counts` may be passed in
    preallocated).  Calls that overlap on one GPU race."""
    import torch
    if not frames.is_cuda:
        raise ValueError("detect_device_static needs a device tensor")
    if frames.dim() == 2:
        frames = frames.unsqueeze(0)
    if not frames.is_contiguous():
        raise ValueError("frames must be contiguous (T, Y, X)")
    T, H, W = frames.shape
    cap = cap_per_frame or max(1024, (H * W) // 16)
    if out is None:
        out = torch.empty((T, cap, 12), dtype=torch.int32, device=frames.device)
    if counts is None:
        counts = torch.empty((T,), dtype=torch.int32, device=frames.device)
    if stream is None:
        stream = torch.cuda.current_stream(frames.device).cuda_stream
    N.check(N.lib().gt_detect_device_static(C.byref(params), H, W, frames.data_ptr(),
                                            _torch_code(frames), T, out.data_ptr(), cap,
                                            counts.data_ptr(), stream))
    return out, counts


def detections_to_numpy(out, counts) -> list:
    """Device detections -> list of per-frame structured arrays sorted by seed."""
    dt = np.dtype(N.DETECTION_FIELDS)
    o = out.cpu().numpy()
    c = counts.cpu().numpy().astype(np.int64)
    res = []
    for t in range(o.shape[0]):
        n = int(c[t])
        if n > o.shape[1]:
            raise N.error(101, f"frame {t}: {n} detections above capacity {o.shape[1]}")
        a = np.ascontiguousarray(o[t, :n]).view(dt).reshape(-1)
        res.append(np.sort(a, order="seed"))
    return res


def _to_dataframe(data: np.ndarray, names, types):
    import pandas as pd
    df = pd.DataFrame(data, columns=names)
    return df.astype({n: ("int64" if t == "int" else "float32") for n, t in zip(names, types)})


def batch(file_or_array, diameter, **kwargs):
    """Trackpy-like detection (and, with search_range, linking) on the GPU.

    Same arguments and defaults as gpu_tracking.batch (generated.py:2-57).
    Returns a pandas DataFrame with columns frame, y, x, mass, [Rg, raw,
    signal, ecc], [raw_mass, raw_bg_median, raw_mass_corrected],
    [particle, duration]."""
    video = _as_video(file_or_array, kwargs.pop("channel", None))
    frames = kwargs.pop("frames", None)
    reset = kwargs.pop("linker_reset_points", None)
    kwargs.pop("tqdm", None)
    learned = kwargs.pop("learned", None)
    learned_conf = kwargs.pop("learned_conf", 0.0)
    learned_tol = kwargs.pop("learned_tol", 0.05)
    learned_scope = kwargs.pop("learned_scope", "moving")
    p = make_params(diameter, **kwargs)
    with Context(p, video.shape[1], video.shape[2]) as ctx:
        if learned is not None:
            ctx.set_learned(learned, learned_conf, learned_tol, learned_scope)
        data, names, types = ctx.execute(video, keys=frames, linker_reset_points=reset)
    return _to_dataframe(data, names, types)


def characterize_points(file_or_array, points_to_characterize, diameter=None, has_frames=None,
                        has_r=None, **kwargs):
    """Characterize given positions (generated.py:59-162)."""
    import pandas as pd
    if isinstance(points_to_characterize, pd.DataFrame):
        cols = ["y", "x"]
        if "frame" in points_to_characterize.columns:
            cols = ["frame"] + cols
            if has_frames is None:
                has_frames = True
        if "r" in points_to_characterize.columns:
            cols = cols + ["r"]
            if has_r is None:
                has_r = True
        points_arr = points_to_characterize[cols].to_numpy().astype("float32")
    else:
        points_arr = np.asarray(points_to_characterize, np.float32)
    has_frames = bool(has_frames)
    has_r = bool(has_r)
    if diameter is None:
        if has_r:
            diameter = 2 * int(points_arr[:, -1].max() + 0.5) + 1
        else:
            raise ValueError("Diameter needs to be specified if the supplied points don't have "
                             "associated radiuses")
    video = _as_video(file_or_array, kwargs.pop("channel", None))
    frames = kwargs.pop("frames", None)
    reset = kwargs.pop("linker_reset_points", None)
    kwargs.pop("tqdm", None)
    p = make_params(diameter, **kwargs)
    # characterize_points_rust forces characterize and disables filter_close
    # (macros lib.rs:671-681); r is reported when the points carry it.
    p.characterize = 1
    p.filter_close = 0
    if has_r:
        p.include_r_in_output = 1
    with Context(p, video.shape[1], video.shape[2], characterize_points=True) as ctx:
        data, names, types = ctx.execute(video, keys=frames, points=points_arr,
                                         points_has_frames=has_frames, points_has_r=has_r,
                                         linker_reset_points=reset)
    return _to_dataframe(data, names, types)


# threshold_rel has no reference counterpart (default off): keep a scale-space
# maximum only where its LoG value exceeds threshold_rel times the frame's
# largest LoG value over every scale (skimage blob_log's threshold_rel)
_LOG_ONLY = dict(n_radii=10, log_spacing=False, overlap_threshold=0.0, threshold_rel=0.0)
_LOG_DEFAULTS = dict({k: v for k, v in _TRACKPY_DEFAULTS.items()
                      if k not in ("maxsize", "separation", "threshold", "invert", "percentile",
                                   "topn", "preprocess", "filter_close")}, **_LOG_ONLY)


def make_log_params(min_radius: float, max_radius: float, **kw) -> N.GtParams:
    """TrackingParams for the LoG style with the Python binding's defaults
    (macros lib.rs:571-629: n_radii 10, linear spacing, overlap_threshold 0,
    the r column on)."""
    unknown = set(kw) - set(_LOG_DEFAULTS)
    if unknown:
        raise TypeError(f"unexpected keyword arguments: {sorted(unknown)}")
    a = dict(_LOG_DEFAULTS)
    a.update({k: v for k, v in kw.items() if v is not None or k in ("search_range", "memory")})
    p = N.default_params()
    p.style = N.GT_STYLE_LOG
    p.min_radius = float(min_radius)
    p.max_radius = float(max_radius)
    p.n_radii = int(a["n_radii"])
    p.log_spacing = int(bool(a["log_spacing"]))
    p.overlap_threshold = float(a["overlap_threshold"])
    p.threshold_rel = float(a["threshold_rel"] or 0.0)
    p.include_r_in_output = 1
    _common(p, a)
    return p


def LoG(file_or_array, min_radius, max_radius, **kwargs):
    """Laplacian-of-Gaussian blob detection (generated.py:164-227; the
    reference wrapper passes the undefined names min_r, max_r, fixed here).

    Returns a pandas DataFrame with columns frame, y, x, mass, r, [Rg, raw,
    signal, ecc], [raw_mass, raw_bg_median, raw_mass_corrected],
    [particle, duration]; rows of a frame in descending r, then mass."""
    video = _as_video(file_or_array, kwargs.pop("channel", None))
    frames = kwargs.pop("frames", None)
    reset = kwargs.pop("linker_reset_points", None)
    kwargs.pop("tqdm", None)
    p = make_log_params(min_radius, max_radius, **kwargs)
    with Context(p, video.shape[1], video.shape[2]) as ctx:
        data, names, types = ctx.execute(video, keys=frames, linker_reset_points=reset)
    return _to_dataframe(data, names, types)


def load(path, frames=None, channel=None) -> np.ndarray:
    """A video file as a (T, Y, X) float32 array (load_rust,
    gpu_tracking_python/src/expansion.rs:758-800); `frames` selects and
    orders frames."""
    with VideoFile(path, channel) as f:
        return f.read(frames).astype(np.float32, copy=False)


def mean_from_file(path, channel=None) -> np.ndarray:
    """The mean frame of a video file (mean_from_file_rust,
    gpu_tracking_python/src/expansion.rs:895-905)."""
    with VideoFile(path, channel) as f:
        return f.mean()


def parse_ets(path) -> dict:
    """Every channel of an ETS file as {channel: (T, Y, X) uint16 array}
    (gpu_tracking_python/src/expansion.rs:849-867)."""
    with VideoFile(path) as f0:
        chans = f0.channels()
    out = {}
    for ch in chans:
        with VideoFile(path, ch) as f:
            out[ch] = f.read()
    return out


def parse_ets_with_keys(path, keys, channel=None) -> np.ndarray:
    """Frames `keys` of one ETS channel as a uint16 array
    (gpu_tracking_python/src/expansion.rs:870-892)."""
    with VideoFile(path, channel) as f:
        return f.read(keys)


def to_tiff(path, video) -> None:
    """Writes a (T, Y, X) video as uncompressed 16-bit grey TIFF pages,
    values cast as `f32 as u16` (gpu_tracking_python/src/expansion.rs:
    925-950)."""
    arr = np.ascontiguousarray(np.asarray(video, np.float32))
    if arr.ndim == 2:
        arr = arr[None]
    if arr.ndim != 3:
        raise ValueError(f"expected a (T, Y, X) array, got shape {arr.shape}")
    N.check(N.lib().gt_write_tiff(os.fsencode(path), arr.ctypes.data, arr.shape[0],
                                  arr.shape[1], arr.shape[2]))


def link(to_link, search_range, memory=0):
    """Links detections into trajectories (lib.py:74-106): to_link is a
    DataFrame with columns frame, y, x (sorted by frame) or an (N, 3) array
    of the same.  Returns the DataFrame with a "particle" column added, or
    the array of particle ids.  (The reference wrapper calls the
    non-existent ndarray.as_type; fixed here.)"""
    import pandas as pd
    if isinstance(to_link, pd.DataFrame):
        arr = to_link[["frame", "y", "x"]].to_numpy()
    else:
        arr = to_link
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise ValueError("expected an (N, 3) array of frame, y, x")
    ids = np.zeros(arr.shape[0], np.uint64)
    N.check(N.lib().gt_link(arr.ctypes.data, arr.shape[0], float(search_range),
                            int(memory or 0), ids.ctypes.data))
    if isinstance(to_link, pd.DataFrame):
        out = to_link.copy()
        out["particle"] = ids
        return out
    return ids


def _link_rows(x, name):
    import pandas as pd
    arr = x[["frame", "y", "x"]].to_numpy() if isinstance(x, pd.DataFrame) else x
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise ValueError(f"{name}: expected an (N, 3) array of frame, y, x")
    return arr


def connect(to_link1, to_link2, search_range, merge=True):
    """Links two sets of detections frame by frame (lib.py:108-158), e.g. two
    channels, or detections against ground truth: each input is a DataFrame
    with columns frame, y, x (sorted by frame) or an (N, 3) array of the same.
    Points of set 2 linked to a point of set 1 share its "connect particle"
    id; the others get ids of their own.  With DataFrames, merge=True returns
    the outer merge of the two on "connect particle" (columns of set 1
    suffixed _x, of set 2 _y; NaN where a point has no partner) and
    merge=False both DataFrames with the column added.  With arrays, the
    two id arrays are returned (the reference would call DataFrame.merge on
    them)."""
    import pandas as pd
    a1 = _link_rows(to_link1, "to_link1")
    a2 = _link_rows(to_link2, "to_link2")
    ids1 = np.zeros(a1.shape[0], np.uint64)
    ids2 = np.zeros(a2.shape[0], np.uint64)
    N.check(N.lib().gt_connect(a1.ctypes.data, a1.shape[0], a2.ctypes.data, a2.shape[0],
                               float(search_range), ids1.ctypes.data, ids2.ctypes.data))
    frames = [isinstance(x, pd.DataFrame) for x in (to_link1, to_link2)]
    if not all(frames):
        out1 = to_link1.assign(**{"connect particle": ids1}) if frames[0] else ids1
        out2 = to_link2.assign(**{"connect particle": ids2}) if frames[1] else ids2
        return out1, out2
    out1 = to_link1.copy()
    out1["connect particle"] = ids1
    out2 = to_link2.copy()
    out2["connect particle"] = ids2
    if not merge:
        return out1, out2
    return out1.merge(out2, how="outer", on="connect particle")


def unique_particle_ids(df, column="particle"):
    """Makes the ids in `column` unique across runs by appending one random
    UUID to every id of this DataFrame, as strings, in place (lib.py:251-253;
    the reference calls the non-existent Series.as_type and adds a UUID
    object to strings - fixed here).  Returns the DataFrame."""
    import uuid
    tag = str(uuid.uuid4())
    df[column] = df[column].astype(str) + tag
    return df


class Batch:
    """Batched detection of frames of different sizes (u8/u16, Trackpy style):
    every launch runs one kernel per stage over a table of frames of any
    sizes, replayed from a HIP graph; results equal Context.detect_device's
    bit for bit.  No counterpart in the reference (one GpuState per frame
    size)."""

    def __init__(self, params: N.GtParams, max_frame_pixels: int,
                 max_launch_pixels: int = 1 << 28, max_slots: int = 256):
        self._lib = N.lib()
        self._h = C.c_void_p()
        N.check(self._lib.gt_batch_create(C.byref(params), int(max_frame_pixels),
                                          int(max_launch_pixels), int(max_slots),
                                          C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            self._lib.gt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def _descs(frames, frame_ids):
        """gt_frame_desc array for `frames`: 2-D frames, or 3-D videos whose
        frames join the list in order (built without a per-frame loop)."""
        import torch
        ptrs, rows, cols = [], [], []
        code = None
        for f in frames:
            if not f.is_cuda or f.dim() not in (2, 3) or not f.is_contiguous():
                raise ValueError("frames must be contiguous 2-D (or 3-D video) device tensors")
            c = N.GT_U16 if f.dtype in (torch.int16, torch.uint16) else (
                N.GT_U8 if f.dtype == torch.uint8 else None)
            if c is None:
                raise TypeError("batched frames must be u8 or u16 (int16 tensors read as u16)")
            if code is not None and c != code:
                raise TypeError("all frames of a batch must have one pixel type")
            code = c
            if f.dim() == 2:
                ptrs.append(np.array([f.data_ptr()], dtype=np.uint64))
                rows.append(np.array([f.shape[0]], dtype=np.uint32))
                cols.append(np.array([f.shape[1]], dtype=np.uint32))
            else:
                t, h, w = f.shape
                step = h * w * f.element_size()
                ptrs.append(np.uint64(f.data_ptr()) + np.arange(t, dtype=np.uint64) * np.uint64(step))
                rows.append(np.full(t, h, dtype=np.uint32))
                cols.append(np.full(t, w, dtype=np.uint32))
        n = int(sum(len(x) for x in ptrs))
        arr = np.zeros(n, dtype=_FRAME_DESC_DTYPE)
        if n:
            arr["data"] = np.concatenate(ptrs)
            arr["nrows"] = np.concatenate(rows)
            arr["ncols"] = np.concatenate(cols)
            arr["frame"] = (np.asarray(frame_ids, dtype=np.uint32) if frame_ids is not None
                            else np.arange(n, dtype=np.uint32))
        return arr, code

    def detect(self, frames, frame_ids=None, out=None, counts=None,
               cap_per_frame: int = 8192, stream=None):
        """frames: list of 2-D device tensors (frames) or 3-D ones (videos,
        whose frames join the list in order).  Returns (rows (n, cap, 12)
        int32 words, counts (n,) int32); rows of the i-th frame at out[i]."""
        import torch
        descs, code = self._descs(frames, frame_ids)
        n = len(descs)
        dev = frames[0].device
        if out is None:
            out = torch.empty((n, cap_per_frame, 12), dtype=torch.int32, device=dev)
        if counts is None:
            counts = torch.zeros((n,), dtype=torch.int32, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        N.check(self._lib.gt_batch_detect(self._h, descs.ctypes.data, n, code, out.data_ptr(),
                                          out.shape[1], counts.data_ptr(), stream))
        return out, counts

    def tune(self, frames, cap_per_frame: int = 8192) -> float:
        """Autotunes the schedule (pixels per launch, slow-walk grid, lanes) on
        sample frames; returns the best time (ms) of one pass over them."""
        import torch
        descs, code = self._descs(frames, None)
        n = len(descs)
        out = torch.empty((n, cap_per_frame, 12), dtype=torch.int32, device=frames[0].device)
        counts = torch.zeros((n,), dtype=torch.int32, device=frames[0].device)
        torch.cuda.synchronize()
        ms = C.c_float()
        N.check(self._lib.gt_batch_tune(self._h, descs.ctypes.data, n, code, out.data_ptr(),
                                        cap_per_frame, counts.data_ptr(), C.byref(ms)))
        return float(ms.value)

    def schedule(self) -> dict:
        """The schedule in use and the last call's grid
        (gt_batch_get_schedule_info)."""
        info = _BatchScheduleInfo()
        N.check(self._lib.gt_batch_get_schedule_info(self._h, C.byref(info)))
        return {"pixels_per_launch": info.pixels_per_launch,
                "walk_blocks": info.walk_blocks,
                "max_frames_per_launch": info.max_frames_per_launch,
                "max_pixels_per_launch": info.max_pixels_per_launch,
                "launches_last_call": info.launches, "lanes": info.lanes,
                "tile_shape": (info.tile_rows, info.tile_cols),
                "tile_blocks_last_call": info.tile_blocks,
                "frames_per_launch_last_call": info.launch_slots,
                "max_size_runs_per_launch_last_call": info.max_segments,
                "graph": bool(info.graph)}

    def set_graph(self, enable: bool) -> None:
        N.check(self._lib.gt_batch_set_graph(self._h, int(bool(enable))))

    def set_lanes(self, lanes: int) -> None:
        """Streams (1..4) that consecutive launches are dealt to."""
        N.check(self._lib.gt_batch_set_lanes(self._h, int(lanes)))


# gt_frame_desc (include/gtrack.h) as a numpy record
_FRAME_DESC_DTYPE = np.dtype([("data", "<u8"), ("nrows", "<u4"), ("ncols", "<u4"),
                              ("frame", "<u4"), ("pad", "<u4")])


class _BatchScheduleInfo(C.Structure):
    # gt_batch_schedule_info (include/gtrack.h)
    _fields_ = [("pixels_per_launch", C.c_uint64), ("max_pixels_per_launch", C.c_uint64),
                ("walk_blocks", C.c_uint32), ("max_frames_per_launch", C.c_uint32),
                ("lanes", C.c_uint32), ("tile_rows", C.c_uint32), ("tile_cols", C.c_uint32),
                ("launches", C.c_uint32), ("tile_blocks", C.c_uint32),
                ("launch_slots", C.c_uint32), ("max_segments", C.c_uint32),
                ("graph", C.c_int32)]


class CentroidNet:
    """Learned centroid proposal (no counterpart in the reference): a small
    CNN that predicts where the centroid walk from a candidate pixel ends.
    Its float32 parameters (torch layouts w1 b1 w2 b2 w3 b3 wfc bfc, 14147
    values), Adam state and bf16 working copies live on the GPU; training
    runs the library's forward/backward kernels."""

    NPARAM = N.CNN_NPARAM

    def __init__(self, seed: int = 0, params=None):
        self._lib = N.lib()
        h = C.c_void_p()
        N.check(self._lib.gt_cnn_create(int(seed), C.byref(h)))
        self._h = h
        if params is not None:
            self.params = params

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.gt_cnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.NPARAM, np.float32)
        N.check(self._lib.gt_cnn_get_params(self._h, out.ctypes.data))
        return out

    @params.setter
    def params(self, p) -> None:
        a = np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1))
        if a.shape != (self.NPARAM,):
            raise ValueError(f"expected {self.NPARAM} parameters, got {a.size}")
        N.check(self._lib.gt_cnn_set_params(self._h, a.ctypes.data))

    @property
    def adam_steps(self) -> int:
        v = C.c_uint64(0)
        N.check(self._lib.gt_cnn_adam_steps(self._h, C.byref(v)))
        return v.value

    def save(self, path: str) -> None:
        np.save(path, self.params)

    @classmethod
    def load(cls, path: str) -> "CentroidNet":
        return cls(params=np.load(path, allow_pickle=False))

    @staticmethod
    def _check_samples(patches, targets=None):
        import torch
        if not patches.is_cuda or patches.dtype != torch.bfloat16 or patches.dim() != 2 \
                or patches.shape[1] != N.CNN_PATCH_PITCH or not patches.is_contiguous():
            raise ValueError("patches must be a contiguous CUDA bfloat16 tensor (n, 128)")
        if targets is not None and (targets.dtype != torch.float32 or targets.shape !=
                                    (patches.shape[0], 4) or not targets.is_contiguous()):
            raise ValueError("targets must be a contiguous float32 tensor (n, 4)")

    def forward(self, patches, stream=None):
        """(n, 4) float32 device tensor of (dy, dx, logit, 0)."""
        import torch
        self._check_samples(patches)
        out = torch.empty((patches.shape[0], 4), dtype=torch.float32, device=patches.device)
        if stream is None:
            stream = torch.cuda.current_stream(patches.device).cuda_stream
        N.check(self._lib.gt_cnn_forward(self._h, patches.data_ptr(), patches.shape[0],
                                         out.data_ptr(), stream))
        return out

    def _idx(self, patches, idx):
        import torch
        if idx is None:
            idx = torch.arange(patches.shape[0], dtype=torch.int32, device=patches.device)
        idx = idx.to(device=patches.device, dtype=torch.int32).contiguous()
        if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= patches.shape[0]):
            raise IndexError("sample index out of range")
        torch.cuda.synchronize(patches.device)
        return idx

    def gradient(self, patches, targets, idx=None):
        """Gradient (numpy float32 (14147,)) and loss of one batch."""
        self._check_samples(patches, targets)
        idx = self._idx(patches, idx)
        g = np.empty(self.NPARAM, np.float32)
        L = C.c_float(0)
        N.check(self._lib.gt_cnn_gradient(self._h, patches.data_ptr(), targets.data_ptr(),
                                          idx.data_ptr(), idx.numel(), g.ctypes.data,
                                          C.byref(L)))
        return g, L.value

    def step(self, patches, targets, idx=None, lr: float = 1e-3) -> float:
        """One Adam step on a batch; returns the batch loss before it."""
        self._check_samples(patches, targets)
        idx = self._idx(patches, idx)
        L = C.c_float(0)
        N.check(self._lib.gt_cnn_step(self._h, patches.data_ptr(), targets.data_ptr(),
                                      idx.data_ptr(), idx.numel(), float(lr), C.byref(L)))
        return L.value

    def train(self, patches, targets, epochs: int = 1, batch: int = 4096, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        """Adam over shuffled batches; returns the mean batch loss per epoch."""
        import torch
        self._check_samples(patches, targets)
        torch.cuda.synchronize(patches.device)
        losses = np.zeros(int(epochs), np.float32)
        N.check(self._lib.gt_cnn_train(self._h, patches.data_ptr(), targets.data_ptr(),
                                       patches.shape[0], int(epochs), int(batch), float(lr),
                                       int(seed), losses.ctypes.data))
        return losses
