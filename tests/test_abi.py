"""C ABI (include/gtrack.h, libgtrack.so) checks that need no GPU: the
library loads, exports every declared entry point, and its host-side logic
(defaults, column names, filter_close, doughnut photometry) matches the
reference semantics and the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gpu_tracking_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gtrack.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gt_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(N.EXPORTS) == declared
    assert L.gt_abi_version() == 2


def test_struct_layouts():
    assert C.sizeof(N.GtResultRow) == 40      # ResultRow: 10 x f32
    assert C.sizeof(N.GtDetection) == 48
    assert np.dtype(N.DETECTION_FIELDS).itemsize == 48


def test_params_default_matches_tracking_params_default():
    # TrackingParams::default() (src/gpu_setup.rs:61-97)
    p = N.default_params()
    assert (p.style, p.diameter, p.separation, p.filter_close) == (0, 9, 10, 1)
    assert (p.characterize, p.doughnut_correction, p.has_search_range) == (0, 0, 0)
    assert p.max_iterations == 10 and p.minmass == 0.0 and p.truncate_preprocessed == 1
    assert abs(p.shift_threshold - 0.6) < 1e-7 and p.noise_size == 1.0


def test_column_names():
    # column_names (src/execute_gpu.rs:482-519)
    p = N.default_params()
    assert N.column_names(p)[0] == ["frame", "y", "x", "mass"]
    p.characterize = 1
    p.doughnut_correction = 1
    p.include_r_in_output = 1
    p.has_search_range = 1
    names, types = N.column_names(p)
    assert names == ["frame", "y", "x", "mass", "r", "Rg", "raw", "signal", "ecc", "raw_mass",
                     "raw_bg_median", "raw_mass_corrected", "particle", "duration"]
    assert types[0] == "int" and types[-1] == "int" and types[-2] == "int"
    assert set(types[1:-2]) == {"float"}


def test_python_defaults():
    # the Python binding's defaults (gpu_tracking_macros/src/lib.rs:187-569)
    from gpu_tracking_amd import make_params
    p = make_params(9)
    assert p.separation == 10 and p.characterize == 1 and p.doughnut_correction == 1
    assert p.filter_close == 1 and p.has_search_range == 0 and p.has_gap_radius == 0
    p = make_params(7, separation=5, bg_radius=8.0, minmass=100)
    assert p.separation == 5 and p.has_bg_radius == 1 and p.has_gap_radius == 1
    assert p.gap_radius == 0.0 and p.minmass == 100.0


def _rows(oracle, n, seed):
    rng = np.random.default_rng(seed)
    rows = np.zeros(n, oracle.ROW_DTYPE)
    rows["x"] = rng.uniform(0, 80, n)
    rows["y"] = rng.uniform(0, 80, n)
    rows["mass"] = rng.integers(1, 30, n)
    return rows


def _as_result_rows(rows):
    a = np.zeros(len(rows), np.dtype(N.DETECTION_FIELDS[:10]))
    for f in a.dtype.names:
        a[f] = rows[f]
    return a


@pytest.mark.parametrize("seed,n,sep", [(0, 400, 5.0), (1, 50, 9.0), (2, 1000, 3.0), (3, 1, 4.0)])
def test_filter_close_matches_oracle(oracle, seed, n, sep):
    rows = _rows(oracle, n, seed)
    rows["x"][::17] = np.nan  # NaN positions never filter and are never filtered
    ref = oracle.filter_close(rows, sep)
    rr = _as_result_rows(rows)
    keep = np.zeros(n, np.uint8)
    N.check(N.lib().gt_filter_close(rr.ctypes.data, n, sep, keep.ctypes.data))
    assert np.array_equal(keep.astype(bool), ref)


def test_doughnut_matches_oracle(oracle):
    rng = np.random.default_rng(4)
    raw = rng.integers(0, 800, (50, 60)).astype(np.float32)
    rows = _rows(oracle, 60, 5)
    rows["x"] = rng.uniform(-2, 52, 60)
    rows["y"] = rng.uniform(-2, 62, 60)
    rows["x"][0], rows["y"][0] = 20.5, -0.5   # halfway cases round away from zero
    for radius, bg, gap in [(4.5, 9.0, 0.0), (3.5, 10.0, 1.5)]:
        ref = oracle.doughnut(raw, rows, radius, bg, gap)
        rr = _as_result_rows(rows)
        out = np.zeros((60, 3), np.float32)
        N.check(N.lib().gt_doughnut(raw.ctypes.data, 50, 60, rr.ctypes.data, 60, radius, bg, gap,
                                    out.ctypes.data))
        assert np.array_equal(out, ref)


def test_context_without_device_reports_adapter_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    p = N.default_params()
    h = C.c_void_p()
    st = N.lib().gt_context_create(C.byref(p), 64, 64, 0, C.byref(h))
    assert st == 1  # GpuAdapterError
    assert b"no HIP device" in N.lib().gt_last_error()


def test_invalid_parameters_rejected():
    p = N.default_params()
    p.diameter = 0
    h = C.c_void_p()
    assert N.lib().gt_context_create(C.byref(p), 64, 64, 0, C.byref(h)) == 100
    p = N.default_params()
    p.has_smoothing_size, p.smoothing_size = 1, 41
    assert N.lib().gt_context_create(C.byref(p), 64, 64, 0, C.byref(h)) == 100
    assert b"smoothing_size" in N.lib().gt_last_error()


@pytest.mark.parametrize("seed,n,thr", [(0, 500, 0.0), (1, 300, 0.3), (2, 40, 0.0), (3, 1, 0.0),
                                        (4, 200, 1.0)])
def test_prune_blobs_matches_oracle(oracle, seed, n, thr):
    rng = np.random.default_rng(seed)
    rows = np.zeros(n, oracle.ROW_DTYPE)
    rows["x"] = rng.uniform(0, 80, n)
    rows["y"] = rng.uniform(0, 80, n)
    rows["r"] = rng.choice(np.float32([1.5, 1.75, 2.04, 3.0, 6.0]), n)
    rows["mass"] = rng.integers(0, 4, n)             # ties in r and mass
    rows["seed"] = np.arange(n)
    if n > 20:
        rows["x"][::19] = np.nan                      # zero-mass walks: never neighbours
    ref = oracle.prune_blobs(rows, thr)
    rr = _as_result_rows(rows)
    nout = C.c_uint64()
    N.check(N.lib().gt_prune_blobs(rr.ctypes.data, n, thr, C.byref(nout)))
    got = rr[:nout.value]
    assert nout.value == len(ref)
    for f in ("x", "y", "r", "mass"):
        assert np.array_equal(got[f].view(np.uint32), ref[f].view(np.uint32)), f


def test_log_parameters_validated():
    import gpu_tracking_amd as gt
    h = C.c_void_p()
    for kw in (dict(n_radii=0), dict(n_radii=65)):
        p = gt.make_log_params(1.5, 6.0, **kw)
        assert N.lib().gt_context_create(C.byref(p), 64, 64, 0, C.byref(h)) == 100
    p = gt.make_log_params(0.0, 6.0)
    assert N.lib().gt_context_create(C.byref(p), 64, 64, 0, C.byref(h)) == 100
    p = gt.make_log_params(1.5, 40.0, smoothing_size=9)     # sigma 28: taps beyond 48
    assert N.lib().gt_context_create(C.byref(p), 64, 64, 0, C.byref(h)) == 100
    assert b"radius" in N.lib().gt_last_error()


def test_frame_desc_record_matches_header_struct():
    # the numpy record the batch wrapper fills is laid out as gt_frame_desc
    import ctypes as C
    from gpu_tracking_amd import api
    dt = api._FRAME_DESC_DTYPE
    assert dt.itemsize == C.sizeof(N.GtFrameDesc) == 24
    for name, _ in N.GtFrameDesc._fields_:
        assert dt.fields[name][1] == getattr(N.GtFrameDesc, name).offset, name
