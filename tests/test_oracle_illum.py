"""The oracle's illumination correction (or_illum_taps, or_illum_smooth,
or_frames_mean, or_illum_divide) pinned on the CPU: closed forms (constant
frames are reproduced exactly, including where taps leave the frame, which
pins the in-frame-only norm), scipy's 1-D correlation in float64 on
non-square frames (pins both passes and their frame bounds), the taps against numpy's
float64 exp, the sequential f32 mean against a float32 loop that is not
numpy's pairwise sum, and power-of-two scale invariance of the per-frame
correction."""
import numpy as np
import pytest

import oracle

scipy_ndimage = pytest.importorskip("scipy.ndimage")


@pytest.mark.parametrize("sigma", [0.5, 1.0, 2.5, 30.0])
def test_taps(sigma):
    w = oracle.illum_taps(sigma)
    rint = int(4 * sigma + 0.5)
    assert w.size == 2 * rint + 1 and w[rint] == 1.0
    assert np.array_equal(w, w[::-1])
    i = np.arange(-rint, rint + 1, dtype=np.float64)
    ref = np.exp(-(i * i) / (2.0 * np.float64(np.float32(sigma)) ** 2))
    assert np.allclose(w, ref, rtol=3e-7, atol=1e-45)


@pytest.mark.parametrize("shape,sigma", [((20, 30), 3.0), ((7, 50), 2.0), ((40, 9), 30.0),
                                         ((1, 13), 1.5)])
@pytest.mark.parametrize("c", [1.0, 8.0, 0.25])
def test_constant_frames_are_kept_exactly(shape, sigma, c):
    # sum of fmaf(w, c, .) over in-frame taps equals c * (sum of those w)
    # exactly for c a power of two: the norm must skip the same taps
    f = np.full(shape, c, np.float32)
    out = oracle.illum_smooth(f, sigma)
    assert np.all(out == np.float32(c))


def _scipy_smooth(f, sigma):
    w = oracle.illum_taps(sigma).astype(np.float64)
    x = f.astype(np.float64)
    one = np.ones_like(x)
    a = scipy_ndimage.correlate1d(x, w, axis=0, mode="constant") / \
        scipy_ndimage.correlate1d(one, w, axis=0, mode="constant")
    return scipy_ndimage.correlate1d(a, w, axis=1, mode="constant") / \
        scipy_ndimage.correlate1d(one, w, axis=1, mode="constant")


@pytest.mark.parametrize("shape,sigma", [((33, 57), 2.0), ((64, 24), 5.0), ((120, 80), 30.0)])
def test_matches_scipy_correlation(shape, sigma):
    rng = np.random.default_rng(sum(shape))
    # non-square, with structure along both axes: a pass bounded by the
    # wrong dimension or a missing pass would show
    r, c = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    f = (100 + 50 * np.sin(r / 3.0) + 0.2 * c * c + rng.normal(0, 5, shape)).astype(np.float32)
    got = oracle.illum_smooth(f, sigma).astype(np.float64)
    ref = _scipy_smooth(f, sigma)
    assert np.allclose(got, ref, rtol=2e-6, atol=0)
    # one pass only (a dropped pass) is far off
    assert not np.allclose(got, scipy_ndimage.correlate1d(f.astype(np.float64),
                                                          oracle.illum_taps(sigma).astype(np.float64),
                                                          axis=0, mode="constant"), rtol=1e-3)


def test_frames_mean_is_sequential_f32():
    rng = np.random.default_rng(4)
    fr = rng.integers(60000, 65535, (500, 5, 6)).astype(np.float32)
    acc = np.zeros((5, 6), np.float32)
    for t in range(fr.shape[0]):
        acc = (acc + fr[t]).astype(np.float32)
    want = (acc / np.float32(fr.shape[0])).astype(np.float32)
    got = oracle.frames_mean(fr)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, fr.astype(np.float64).mean(0).astype(np.float32))


def test_correction_divides_by_the_smoothed_mean():
    rng = np.random.default_rng(6)
    v = rng.integers(100, 4000, (4, 30, 40)).astype(np.float32)
    out = oracle.illum_correct(v, 4.0)
    prof = oracle.illum_smooth(oracle.frames_mean(v), 4.0)
    assert np.array_equal(out, (v / prof[None]).astype(np.float32))


def test_per_frame_correction_is_scale_invariant():
    # a frame scaled by 2^k smooths to the scaled profile exactly, so the
    # per-frame corrections coincide; the whole-video correction does not
    rng = np.random.default_rng(7)
    base = rng.integers(100, 4000, (30, 40)).astype(np.float32)
    v = np.stack([base, base * 4, base / 8]).astype(np.float32)
    out = oracle.illum_correct(v, 3.0, per_frame=True)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    whole = oracle.illum_correct(v, 3.0)
    assert not np.array_equal(whole[0], whole[1])
