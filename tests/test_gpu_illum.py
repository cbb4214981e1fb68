"""Illumination correction on the GPU (the whole video's smoothed mean frame,
or each frame's own smoothed copy, divides the frames before detection):
detections on the corrected frames bit for bit against the oracle's
correction followed by its detection, through the device path and through
batch() (whose doughnut reads the uncorrected frames, as the reference's
host doughnut does), and through a video file."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


def _video(T, seed, H, W):
    from gpu_tracking_amd import synth
    return synth.video_numpy(synth.spec("full", nrows=H, ncols=W,
                                        n_particles=max(8, H * W // 1000)), T, seed)


def _dev(torch, v):
    return torch.from_numpy(v.view(np.int16)).cuda()


def _assert_rows_equal(got, ref, label):
    assert len(got) == len(ref), f"{label}: {len(got)} rows vs oracle {len(ref)}"
    assert np.array_equal(got["seed"], ref["seed"]), label
    for f in FIELDS:
        a, b = got[f], ref[f]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert np.all(same), f"{label}: field {f} differs in {int((~same).sum())} rows"


@pytest.mark.parametrize("sigma,per_frame", [(4.0, False), (4.0, True), (30.0, False)])
def test_device_detection_on_corrected_frames(torch_dev, sigma, per_frame):
    import oracle
    import gpu_tracking_amd as gt
    H, W = 160, 224
    v = _video(5, 13, H, W)
    p = gt.make_params(9, separation=7, minmass=2.0, characterize=True,
                       illumination_sigma=sigma, illumination_correction_per_frame=per_frame)
    with gt.Context(p, H, W) as ctx:
        out, counts = ctx.detect_device(_dev(torch_dev, v), cap_per_frame=(H * W) // 4)
        torch_dev.cuda.synchronize()
        rows = gt.detections_to_numpy(out, counts)
    corr = oracle.illum_correct(v, sigma, per_frame=per_frame)
    op = oracle.derive_trackpy(H, W, diameter=9, separation=7, minmass=2.0)
    total = 0
    for t in range(v.shape[0]):
        ref, _ = oracle.detect_frame(corr[t], op, characterize=True)
        _assert_rows_equal(rows[t], np.sort(ref, order="seed"), f"frame {t}")
        total += len(ref)
    assert total > 20


def _expected_batch(oracle, v, sigma, minmass):
    T, H, W = v.shape
    corr = oracle.illum_correct(v, sigma)
    op = oracle.derive_trackpy(H, W, diameter=9, separation=10, minmass=minmass)
    out = []
    for t in range(T):
        rows, _ = oracle.detect_frame(corr[t], op, characterize=True)
        rows = np.sort(rows, order="seed")
        rows = rows[oracle.filter_close(rows, 10.0)]
        d = oracle.doughnut(v[t], rows, 4.5, 9.0, 0.0)
        cols = [np.full(len(rows), t, np.float32), rows["x"], rows["y"], rows["mass"],
                rows["Rg"], rows["raw_mass"], rows["signal"], rows["ecc"],
                d[:, 0], d[:, 1], d[:, 2]]
        out.append(np.stack(cols, 1).astype(np.float32))
    return np.concatenate(out, 0)


def test_batch_with_correct_illumination(torch_dev, tmp_path):
    import oracle
    import gpu_tracking_amd as gt
    v = _video(4, 17, 128, 160)
    df = gt.batch(v, 9, minmass=2.0, correct_illumination=True, illumination_sigma=6.0)
    got = df.to_numpy(dtype=np.float64).astype(np.float32)
    ref = _expected_batch(oracle, v, 6.0, 2.0)
    assert got.shape == ref.shape and len(ref) > 10
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # from a file: the profile is the mean of every frame of the file
    p = str(tmp_path / "i.tif")
    gt.to_tiff(p, v.astype(np.float32))
    df2 = gt.batch(p, 9, minmass=2.0, correct_illumination=True, illumination_sigma=6.0)
    got2 = df2.to_numpy(dtype=np.float64).astype(np.float32)
    assert np.array_equal(got2.view(np.uint32), ref.view(np.uint32))


def test_batch_context_rejects_illumination(torch_dev):
    import gpu_tracking_amd as gt
    p = gt.make_params(9, illumination_sigma=5.0)
    with pytest.raises(gt.GtError):
        gt.Batch(p, 1 << 20)
