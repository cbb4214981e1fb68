"""One frame at a time through the C ABI: gt_submit_frame / gt_fetch_frame
(the reference's submit_work / get_work, src/execute_gpu.rs:95-105,424-467)
against gt_execute over the same video, bit for bit, and against the oracle."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "mass", "Rg", "raw_mass", "signal", "ecc"]


@pytest.fixture(scope="module")
def gt(gpu_ok):
    import gpu_tracking_amd
    return gpu_tracking_amd


def _video(T, seed, H=200, W=232, n=30):
    from gpu_tracking_amd import synth
    return synth.video_numpy(synth.spec("small", nrows=H, ncols=W, n_particles=n), T, seed)


PRM = dict(minmass=300.0, characterize=True, filter_close=False, doughnut_correction=False)


def _execute_rows(ctx, v):
    data, names, _ = ctx.execute(v)
    assert names == ["frame", "y", "x", "mass", "Rg", "raw", "signal", "ecc"]
    return data


def _as_matrix(idx, rows):
    return np.stack([np.full(len(rows), idx, np.float32), rows["x"], rows["y"], rows["mass"],
                     rows["Rg"], rows["raw_mass"], rows["signal"], rows["ecc"]], 1) \
        if len(rows) else np.zeros((0, 8), np.float32)


@pytest.mark.parametrize("depth,dtype", [(2, np.uint16), (3, np.uint16), (1, np.float32),
                                         (2, np.uint8)])
def test_streamed_frames_equal_gt_execute(gt, oracle, depth, dtype):
    v = _video(12, 41)
    if dtype == np.uint8:
        v = (v // 64).astype(np.uint8)
    else:
        v = v.astype(dtype)
    prm = dict(PRM, minmass=20.0 if dtype == np.uint8 else 300.0)
    print(f"[stream] depth={depth} dtype={np.dtype(dtype).name}", flush=True)
    with gt.Context(gt.make_params(9, **prm), *v.shape[1:]) as ctx:
        got = list(ctx.stream(v, depth=depth))
        want = _execute_rows(ctx, v)
    assert [k for k, _ in got] == list(range(12))
    m = np.concatenate([_as_matrix(k, r) for k, r in got], 0)
    assert m.shape == want.shape and len(m) > 12
    assert np.array_equal(m.view(np.uint32), want.view(np.uint32))
    p = oracle.derive_trackpy(*v.shape[1:], diameter=9, separation=10, minmass=prm["minmass"])
    for t in (0, 11):
        ref, _ = oracle.detect_frame(v[t], p, characterize=True)
        ref = np.sort(ref, order="seed")
        rows = got[t][1]
        assert len(rows) == len(ref)
        for f in FIELDS:
            assert np.array_equal(rows[f].view(np.uint32), ref[f].view(np.uint32)), f


def test_ring_rules_through_the_c_abi(gt):
    from gpu_tracking_amd import _native as N
    L = N.lib()
    v = _video(3, 42)
    with gt.Context(gt.make_params(9, **PRM), *v.shape[1:]) as ctx:
        h = ctx.handle
        idx, n = C.c_uint64(0), C.c_uint32(0)
        row = np.zeros(1, np.float32)
        # nothing in flight
        assert L.gt_fetch_frame(h, C.byref(idx), None, 0, C.byref(n)) == 100
        # wrong pixel count: Error::DimensionMismatch
        f0 = np.ascontiguousarray(v[0])
        assert L.gt_submit_frame(h, f0.ctypes.data, 0, f0.size - 1, 0, None, 0) == 4
        for k in range(2):
            fk = np.ascontiguousarray(v[k])
            assert L.gt_submit_frame(h, fk.ctypes.data, 0, fk.size, 10 + k, None, 0) == 0
        # the ring (depth 2) is full
        f2 = np.ascontiguousarray(v[2])
        assert L.gt_submit_frame(h, f2.ctypes.data, 0, f2.size, 12, None, 0) == 100
        # a device-frame call while frames are in flight is refused
        import torch
        d = torch.from_numpy(v[:1].view(np.int16)).cuda()
        with pytest.raises(ValueError):
            ctx.detect_device(d)
        # too small a buffer: capacity error, the frame stays in flight
        assert L.gt_fetch_frame(h, C.byref(idx), row.ctypes.data, 1, C.byref(n)) == 101
        assert idx.value == 10 and n.value > 1
        need = n.value
        k, rows = ctx.fetch_frame(cap=need)
        assert k == 10 and len(rows) == need
        k, rows = ctx.fetch_frame()
        assert k == 11
        # the submitted host buffer may be reused at once: it was staged
        buf = np.ascontiguousarray(v[2]).copy()
        ctx.submit_frame(buf, 12)
        buf[:] = 0
        k, rows2 = ctx.fetch_frame()
        ref = [r for r in ctx.stream([v[2]])][0][1]
        assert k == 12 and np.array_equal(rows2.view(np.uint32), ref.view(np.uint32))


def test_streamed_characterize_points_equal_gt_execute(gt):
    v = _video(4, 43)
    T, H, W = v.shape
    rng = np.random.default_rng(5)
    pts = {t: np.column_stack([rng.uniform(5, H - 5, 20), rng.uniform(5, W - 5, 20)])
           .astype(np.float32) for t in range(T)}
    p = gt.make_params(9, characterize=True, doughnut_correction=False)
    p.filter_close = 0
    with gt.Context(p, H, W, characterize_points=True) as ctx:
        got = list(ctx.stream(v, positions=lambda k: pts[k]))
        allpts = np.concatenate([np.column_stack([np.full(20, t, np.float32), pts[t]])
                                 for t in range(T)])
        data, names, _ = ctx.execute(v, points=allpts, points_has_frames=True)
    assert names[:4] == ["frame", "y", "x", "mass"]
    m = np.concatenate([_as_matrix(k, r) for k, r in got], 0)
    assert np.array_equal(m.view(np.uint32), data.view(np.uint32))
