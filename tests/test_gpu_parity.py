"""GPU parity: the HIP pipeline (through the C ABI) against the CPU oracle on
the same seeded frames.  Integer/index results (candidate sets, seeds,
counts, raw sums) must be bit-identical; float columns are compared bit for
bit as well, since both sides perform the same IEEE operations in the same
order (taps in order with fmaf, correctly rounded division and sqrt)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


def _video(name, T, seed, **kw):
    from gpu_tracking_amd import synth
    return synth.video_numpy(synth.spec(name, **kw), T, seed)


def _gpu_detect(torch, video, cap=None, **params):
    import gpu_tracking_amd as gt
    p = gt.make_params(params.pop("diameter", 9), **params)
    T, H, W = video.shape
    with gt.Context(p, H, W) as ctx:
        if video.dtype == np.uint16:
            t = torch.from_numpy(video.view(np.int16)).cuda()
        else:
            t = torch.from_numpy(video).cuda()
        out, counts = ctx.detect_device(t, cap_per_frame=cap or (H * W) // 4)
        torch.cuda.synchronize()
        return gt.detections_to_numpy(out, counts)


def _oracle_rows(oracle, frame, **kw):
    H, W = frame.shape
    ch = kw.pop("characterize", True)
    p = oracle.derive_trackpy(H, W, **kw)
    rows, nc = oracle.detect_frame(frame, p, characterize=ch)
    return np.sort(rows, order="seed"), nc


def _assert_rows_equal(got, ref, label):
    assert len(got) == len(ref), f"{label}: {len(got)} rows vs oracle {len(ref)}"
    assert np.array_equal(got["seed"], ref["seed"]), label
    for f in FIELDS:
        a, b = got[f], ref[f]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        if not np.all(same):
            i = int(np.flatnonzero(~same)[0])
            raise AssertionError(f"{label}: field {f} differs at row {i} (seed {ref['seed'][i]}): "
                                 f"{a[i]!r} vs oracle {b[i]!r}; {int((~same).sum())} rows differ")


CASES = [
    # (generator, frames, seed, params)
    ("small", 3, 101, dict(diameter=9, separation=9, minmass=500.0, characterize=True)),
    ("full", 2, 202, dict(diameter=9, separation=7, minmass=400.0, characterize=True)),
]


@pytest.mark.parametrize("name,T,seed,params", CASES)
def test_candidate_set_bit_identical(torch_dev, oracle, name, T, seed, params):
    # max_iterations = 0 and minmass = 0 turn every candidate into a row at
    # its own pixel, so the row seeds are exactly the candidate set
    v = _video(name, T, seed)
    prm = dict(params, max_iterations=0, minmass=0.0, characterize=False)
    got = _gpu_detect(torch_dev, v, **prm)
    for t in range(T):
        H, W = v.shape[1:]
        p = oracle.derive_trackpy(H, W, diameter=prm["diameter"], separation=prm["separation"])
        P = oracle.bandpass(v[t], p)
        cand = oracle.local_max(P, p)
        assert np.array_equal(got[t]["seed"], cand), f"frame {t}"
        assert np.array_equal(got[t]["x"], (cand // W).astype(np.float32))
        assert np.array_equal(got[t]["y"], (cand % W).astype(np.float32))


@pytest.mark.parametrize("name,T,seed,params", CASES)
def test_rows_bit_identical(torch_dev, oracle, name, T, seed, params):
    v = _video(name, T, seed)
    got = _gpu_detect(torch_dev, v, **params)
    for t in range(T):
        kw = dict(params)
        ref, _ = _oracle_rows(oracle, v[t], diameter=kw.pop("diameter"),
                              separation=kw.pop("separation"), minmass=kw.pop("minmass"),
                              characterize=kw.pop("characterize"))
        _assert_rows_equal(got[t], ref, f"{name} frame {t}")


@pytest.mark.parametrize("H,W", [(37, 53), (100, 65), (64, 64), (31, 200), (257, 129)])
def test_ragged_shapes(torch_dev, oracle, H, W):
    v = _video("small", 2, H * 1000 + W, nrows=H, ncols=W, n_particles=max(1, H * W // 1500),
               edge=min(10.0, H / 4, W / 4), min_dist=6.0)
    prm = dict(diameter=7, separation=6, minmass=50.0, characterize=True)
    got = _gpu_detect(torch_dev, v, **prm)
    for t in range(2):
        ref, _ = _oracle_rows(oracle, v[t], **prm)
        _assert_rows_equal(got[t], ref, f"{H}x{W} frame {t}")


@pytest.mark.parametrize("H,W", [(1, 1), (2, 9), (9, 9), (1, 300), (300, 1), (12, 13)])
def test_tiny_frames(torch_dev, oracle, H, W):
    # frames smaller than a tile or than the margin: every stage's bounds
    # handling, and no rows where the margin leaves no pixel
    rng = np.random.default_rng(H * 1000 + W)
    v = rng.integers(90, 110, (2, H, W)).astype(np.uint16)
    v[:, H // 2, W // 2] += 3000
    prm = dict(diameter=9, separation=9, minmass=0.0, characterize=True)
    got = _gpu_detect(torch_dev, v, cap=max(16, H * W), **prm)
    for t in range(2):
        ref, _ = _oracle_rows(oracle, v[t], **prm)
        _assert_rows_equal(got[t], ref, f"{H}x{W} frame {t}")


@pytest.mark.parametrize("dtype", ["f32", "u8"])
def test_other_pixel_types(torch_dev, oracle, dtype):
    v = _video("small", 2, 77, nrows=128, ncols=192, n_particles=20)
    if dtype == "f32":
        v = v.astype(np.float32) * np.float32(0.37)
    else:
        v = np.clip(v // 8, 0, 255).astype(np.uint8)
    prm = dict(diameter=9, separation=9, minmass=20.0, characterize=True)
    got = _gpu_detect(torch_dev, v, **prm)
    for t in range(2):
        ref, _ = _oracle_rows(oracle, v[t].astype(np.float32), **prm)
        _assert_rows_equal(got[t], ref, f"{dtype} frame {t}")


@pytest.mark.parametrize("params", [
    dict(diameter=5, separation=4, minmass=10.0),
    dict(diameter=11, separation=12, smoothing_size=15, noise_size=2.0, minmass=100.0),
    dict(diameter=9, separation=40, minmass=100.0),            # untiled local max (k = 56)
    dict(diameter=9, separation=9, minmass=0.0, truncate_preprocessed=False),
    dict(diameter=21, separation=9, smoothing_size=21, minmass=0.0),  # generic radius paths
])
def test_parameter_variants(torch_dev, oracle, params):
    v = _video("small", 1, 5, nrows=160, ncols=176, n_particles=25)
    prm = dict(params, characterize=True)
    got = _gpu_detect(torch_dev, v, **prm)
    ref, _ = _oracle_rows(oracle, v[0], **prm)
    _assert_rows_equal(got[0], ref, str(params))


def test_constant_and_empty_frames(torch_dev, oracle):
    # a constant frame truncates to zero everywhere: every pixel inside the
    # margin is a (zero-mass) candidate; with minmass > 0 nothing survives
    v = np.full((2, 64, 96), 100, np.uint16)
    got = _gpu_detect(torch_dev, v, diameter=9, separation=9, minmass=1.0, characterize=True)
    assert all(len(g) == 0 for g in got)
    got = _gpu_detect(torch_dev, v, cap=64 * 96, diameter=9, separation=9, minmass=0.0,
                      characterize=False, max_iterations=0)
    m = 4
    assert len(got[0]) == (64 - 2 * m) * (96 - 2 * m)


def test_frame_batch_invariance(torch_dev):
    # a frame gives the same rows whatever its position in the batch
    v = _video("small", 6, 33, nrows=256, ncols=256, n_particles=30)
    prm = dict(diameter=9, separation=9, minmass=300.0, characterize=True)
    a = _gpu_detect(torch_dev, v, **prm)
    b = _gpu_detect(torch_dev, v[::-1].copy(), **prm)
    for t in range(6):
        _assert_rows_equal(a[t], b[5 - t], f"frame {t}")


def test_capacity_overflow_reported(torch_dev):
    v = np.full((1, 64, 64), 7, np.uint16)
    import gpu_tracking_amd as gt
    with pytest.raises(gt.GtError):
        _gpu_detect(torch_dev, v, cap=16, diameter=9, separation=9, minmass=0.0,
                    max_iterations=0)


def test_two_lanes_give_the_same_rows(torch_dev, oracle):
    # chunks alternating between two streams with their own scratch, and the
    # chunk tails on a second stream (overlap).  A context takes up to 512
    # frames per chunk: 1600 frames of 64 x 96 make four chunks, so both lanes
    # and the overlap really run; frames of chunks 1..3 are also checked
    # against the oracle.
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    H, W, T = 64, 96, 1600
    v = synth.video_numpy(synth.spec("full", nrows=H, ncols=W, n_particles=8), T, 91)
    d = torch_dev.from_numpy(v.view(np.int16)).cuda()
    res = []
    for lanes, overlap in ((1, False), (2, False), (1, True)):
        with gt.Context(gt.make_params(9, separation=7, minmass=300.0), H, W) as ctx:
            ctx.set_lanes(lanes)
            ctx.set_overlap(overlap)
            ctx.set_timing(True)
            out, counts = ctx.detect_device(d, cap_per_frame=512)
            torch_dev.cuda.synchronize()
            ms, launches, per_launch = ctx.stage_times()
            assert (launches, per_launch) == (4, 512), (launches, per_launch)
            res.append(gt.detections_to_numpy(out, counts))
    assert sum(len(r) for r in res[0]) > 0
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for t in (600, 1111, 1599):
        ref, _ = _oracle_rows(oracle, v[t], diameter=9, separation=7, minmass=300.0)
        _assert_rows_equal(res[1][t], ref, f"frame {t}")
