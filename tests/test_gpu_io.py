"""Videos read from files (gt_execute_file: frames decoded by reader threads
straight into the pinned staging) give the same DataFrames, bit for bit, as
the same frames passed as arrays (gt_execute), for batch, LoG and
characterize_points, with frame selection and linking."""
import numpy as np
import pytest

from filefmt import ets_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gt(gpu_ok):
    import gpu_tracking_amd
    return gpu_tracking_amd


def _video(T=70, seed=21, H=160, W=192, n=40):
    from gpu_tracking_amd import synth
    return synth.video_numpy(synth.spec("full", nrows=H, ncols=W, n_particles=n), T, seed)


def _same(a, b):
    assert list(a.columns) == list(b.columns)
    x = a.to_numpy(dtype=np.float64).astype(np.float32)
    y = b.to_numpy(dtype=np.float64).astype(np.float32)
    assert x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_batch_from_tiff_equals_array(gt, tmp_path):
    v = _video()
    p = str(tmp_path / "v.tif")
    gt.to_tiff(p, v.astype(np.float32))
    kw = dict(minmass=300, separation=7, search_range=5, memory=2)
    _same(gt.batch(p, 9, **kw), gt.batch(v, 9, **kw))
    keys = [5, 3, 60, 61, 0]
    _same(gt.batch(p, 9, frames=keys, minmass=300), gt.batch(v, 9, frames=keys, minmass=300))


def test_log_and_characterize_from_tiff(gt, tmp_path):
    v = _video(T=6)
    p = str(tmp_path / "l.tif")
    gt.to_tiff(p, v.astype(np.float32))
    _same(gt.LoG(p, 1.5, 4.0, n_radii=5), gt.LoG(v, 1.5, 4.0, n_radii=5))
    det = gt.batch(v, 9, minmass=300)
    _same(gt.characterize_points(p, det, diameter=9), gt.characterize_points(v, det, diameter=9))


def test_batch_from_ets_channel(gt, tmp_path):
    v = _video(T=9, seed=4)
    w = _video(T=5, seed=5)
    b = ets_bytes({0: list(w), 1: list(v)}, (v.shape[2], v.shape[1]))
    p = str(tmp_path / "x.ets")
    with open(p, "wb") as f:
        f.write(b)
    _same(gt.batch(p, 9, channel=1, minmass=300), gt.batch(v, 9, minmass=300))
    _same(gt.batch(p, 9, minmass=300), gt.batch(w, 9, minmass=300))
