"""adaptive_background on the GPU against the oracle (or_detect_frame_adaptive,
or_log_detect_frame_adaptive): the accepted rows, characterized, must be
bit-identical for several pass counts and rule settings, in the Trackpy and
the LoG style (whose rows mark disks of their own radius)."""
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
    with gt.Context(gt.make_params(9, adaptive_background=1, snr=1.0), 128, 160) as ctx:
        with pytest.raises(gt.GtError):
            ctx.set_learned(gt.CentroidNet(seed=1))


@pytest.mark.parametrize("n_iter", [1, 2, 3])
def test_log_adaptive_rows_bit_identical(torch_dev, n_iter):
    import oracle
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    H, W = 160, 176
    v = synth.video_numpy(synth.spec("log", nrows=H, ncols=W, n_particles=50), 2, 60 + n_iter)
    p = gt.make_log_params(1.5, 4.0, n_radii=5, snr=1.5, minmass_snr=0.3,
                           adaptive_background=n_iter, characterize=True)
    with gt.Context(p, H, W) as ctx:
        out, counts = ctx.detect_device(torch_dev.from_numpy(v).cuda(), cap_per_frame=H * W)
        torch_dev.cuda.synchronize()
        rows = gt.detections_to_numpy(out, counts)
    op = oracle.derive_log(H, W, 4.0)
    op.snr, op.minmass_snr, op.rough_snr_factor = 1.5, 0.3, 0.5
    radii = oracle.log_radii(1.5, 4.0, 5, False)
    total = 0
    for t in range(v.shape[0]):
        ref = np.sort(oracle.log_detect_frame_adaptive(v[t], op, radii, n_iter,
                                                       overlap_threshold=1.0), order="seed")
        _assert_rows_equal(rows[t], ref, f"LoG frame {t} n_iter {n_iter}")
        total += len(ref)
    assert total > 0
    # through LoG() (pruning at the default overlap_threshold 0)
    df = gt.LoG(v, 1.5, 4.0, n_radii=5, snr=1.5, minmass_snr=0.3, adaptive_background=n_iter,
                doughnut_correction=False)
    n_ref = sum(len(oracle.log_detect_frame_adaptive(v[t], op, radii, n_iter,
                                                     overlap_threshold=0.0)) for t in range(2))
    assert len(df) == n_ref
