"""adaptive_background on the GPU against the oracle (or_detect_frame_adaptive):
the accepted rows, characterized, must be bit-identical for several pass
counts and rule settings."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


def _assert_rows_equal(got, ref, label):
    assert len(got) == len(ref), f"{label}: {len(got)} rows vs oracle {len(ref)}"
    assert np.array_equal(got["seed"], ref["seed"]), label
    for f in FIELDS:
        a, b = got[f], ref[f]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert np.all(same), f"{label}: field {f} differs in {int((~same).sum())} rows"


@pytest.mark.parametrize("n_iter,kw", [(1, dict(snr=1.0, minmass_snr=0.2)),
                                       (2, dict(snr=1.5, minmass_snr=0.3)),
                                       (3, dict(snr=2.0, minmass_snr=0.5)),
                                       (2, dict(minmass=300.0, snr=2.0))])
def test_adaptive_rows_bit_identical(torch_dev, n_iter, kw):
    import oracle
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    H, W = 224, 256
    v = synth.video_numpy(synth.spec("full", nrows=H, ncols=W, n_particles=70), 3, 40 + n_iter)
    p = gt.make_params(9, separation=7, characterize=True, adaptive_background=n_iter, **kw)
    with gt.Context(p, H, W) as ctx:
        out, counts = ctx.detect_device(torch_dev.from_numpy(v.view(np.int16)).cuda(),
                                        cap_per_frame=(H * W) // 4)
        torch_dev.cuda.synchronize()
        rows = gt.detections_to_numpy(out, counts)
    op = oracle.derive_trackpy(H, W, diameter=9, separation=7, adaptive_background=True,
                               minmass=kw.get("minmass", 0.0), snr=kw.get("snr", 0.0),
                               minmass_snr=kw.get("minmass_snr", 0.0))
    total = 0
    for t in range(v.shape[0]):
        ref = np.sort(oracle.detect_frame_adaptive(v[t].astype(np.float32), op, n_iter),
                      order="seed")
        _assert_rows_equal(rows[t], ref, f"frame {t} n_iter {n_iter} {kw}")
        total += len(ref)
    assert total > 0


def test_adaptive_through_batch_and_rejections(torch_dev):
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    v = synth.video_numpy(synth.spec("full", nrows=128, ncols=160, n_particles=30), 2, 3)
    df = gt.batch(v, 9, separation=7, snr=1.0, minmass_snr=0.2, adaptive_background=2)
    assert len(df) > 0
    with pytest.raises(gt.GtError):
        gt.batch(v, 9, adaptive_background=0)
    with pytest.raises(gt.GtError):
        gt.LoG(v, 1.5, 3.0, adaptive_background=1)
    with gt.Context(gt.make_params(9, adaptive_background=1, snr=1.0), 128, 160) as ctx:
        with pytest.raises(gt.GtError):
            ctx.set_learned(gt.CentroidNet(seed=1))
