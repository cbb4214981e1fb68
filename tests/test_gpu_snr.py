"""The snr / minmass_snr rules on the GPU: the per-frame noise estimate
(eight-way tree std of the raw frame) bit for bit against the oracle, and
detections with the snr candidate threshold and the minmass_snr rule bit for
bit against the oracle given that estimate."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MARKER = np.array([0xFFFFFFFF], np.uint32).view(np.float32)[0]
FIELDS = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


def _video(T, seed, H, W):
    from gpu_tracking_amd import synth
    if min(H, W) < 32:     # too small for the particle generator's edge
        return np.random.default_rng(seed).poisson(120, (T, H, W)).astype(np.uint16)
    return synth.video_numpy(synth.spec("full", nrows=H, ncols=W,
                                        n_particles=max(8, H * W // 1000)), T, seed)


def _dev(torch, v):
    return torch.from_numpy(v.view(np.int16) if v.dtype == np.uint16 else v).cuda()


@pytest.mark.parametrize("H,W", [(512, 512), (384, 384), (100, 77), (1024, 1024), (5, 7)])
def test_frame_std_bit_identical(torch_dev, H, W):
    import oracle
    import gpu_tracking_amd as gt
    v = _video(3, 31 + H, H, W)
    with gt.Context(gt.make_params(9), H, W) as ctx:
        got = ctx.frame_std(_dev(torch_dev, v)).cpu().numpy()
    for t in range(v.shape[0]):
        _, want = oracle.frame_std(v[t].astype(np.float32))
        assert got[t].view(np.uint32) == np.float32(want).view(np.uint32), (t, got[t], want)


def test_frame_std_float_frames_with_markers(torch_dev):
    import oracle
    import gpu_tracking_amd as gt
    rng = np.random.default_rng(5)
    v = rng.normal(100, 7, (2, 96, 130)).astype(np.float32)
    v.reshape(2, -1)[:, rng.choice(96 * 130, 300, replace=False)] = MARKER
    with gt.Context(gt.make_params(9), 96, 130) as ctx:
        got = ctx.frame_std(_dev(torch_dev, v)).cpu().numpy()
    for t in range(2):
        _, want = oracle.frame_std(v[t])
        assert got[t].view(np.uint32) == np.float32(want).view(np.uint32)


def _assert_rows_equal(got, ref, label):
    assert len(got) == len(ref), f"{label}: {len(got)} rows vs oracle {len(ref)}"
    assert np.array_equal(got["seed"], ref["seed"]), label
    for f in FIELDS:
        a, b = got[f], ref[f]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert np.all(same), f"{label}: field {f} differs in {int((~same).sum())} rows"


@pytest.mark.parametrize("kw", [dict(snr=1.5, minmass=300.0),
                                dict(minmass_snr=0.5),
                                dict(snr=0.8, minmass_snr=0.3)])
def test_detection_with_snr_rules(torch_dev, kw):
    import oracle
    import gpu_tracking_amd as gt
    H, W = 256, 320
    v = _video(3, 77, H, W)
    p = gt.make_params(9, separation=7, characterize=True, **kw)
    with gt.Context(p, H, W) as ctx:
        out, counts = ctx.detect_device(_dev(torch_dev, v), cap_per_frame=(H * W) // 4)
        torch_dev.cuda.synchronize()
        rows = gt.detections_to_numpy(out, counts)
    n_plain = n_thr = 0
    for t in range(v.shape[0]):
        _, std = oracle.frame_std(v[t].astype(np.float32))
        op = oracle.derive_trackpy(H, W, diameter=9, separation=7, minmass=kw.get("minmass", 0.0),
                                   snr=kw.get("snr", 0.0), minmass_snr=kw.get("minmass_snr", 0.0),
                                   image_std=std)
        ref, nc = oracle.detect_frame(v[t], op, characterize=True)
        _assert_rows_equal(rows[t], np.sort(ref, order="seed"), f"frame {t} {kw}")
        # the rules act: fewer rows than without them
        op0 = oracle.derive_trackpy(H, W, diameter=9, separation=7,
                                    minmass=kw.get("minmass", 0.0))
        ref0, _ = oracle.detect_frame(v[t], op0, characterize=False)
        n_plain += len(ref0)
        n_thr += len(ref)
    assert n_thr < n_plain


def test_log_with_snr_threshold(torch_dev):
    import oracle
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    H, W = 192, 160
    v = synth.video_numpy(synth.spec("log", nrows=H, ncols=W, n_particles=60), 2, 9)
    p = gt.make_log_params(1.5, 4.0, n_radii=5, snr=2.0, characterize=True)
    with gt.Context(p, H, W) as ctx:
        out, counts = ctx.detect_device(_dev(torch_dev, v), cap_per_frame=H * W)
        torch_dev.cuda.synchronize()
        rows = gt.detections_to_numpy(out, counts)
    radii = oracle.log_radii(1.5, 4.0, 5, False)
    for t in range(2):
        _, std = oracle.frame_std(v[t])
        op = oracle.derive_log(H, W, 4.0)
        op.snr = 2.0
        op.image_std = std
        ref, _ = oracle.log_detect_frame(v[t], op, radii, characterize=True,
                                         overlap_threshold=1.0)
        _assert_rows_equal(rows[t], np.sort(ref, order="seed"), f"LoG frame {t}")


def test_batch_context_rejects_snr(torch_dev):
    import gpu_tracking_amd as gt
    with pytest.raises(gt.GtError):
        gt.Batch(gt.make_params(9, snr=1.0), max_frame_pixels=64 * 64)
