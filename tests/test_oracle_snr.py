"""The oracle's noise estimate (or_frame_std, the snr / minmass_snr rules'
std of the raw frame) pinned on the CPU: closed forms, numpy's float64
statistics, the eight-way tree's summation order on a known answer, the
NaN-marker rule, and an independent vectorised restatement of the shaders."""
import numpy as np
import pytest

import oracle

MARKER = np.array([0xFFFFFFFF], np.uint32).view(np.float32)[0]


def test_constant_frame():
    m, s = oracle.frame_std(np.full((37, 41), 123.0, np.float32))
    assert m == 123.0 and s == 0.0


@pytest.mark.parametrize("shape", [(64, 64), (100, 77), (512, 512), (384, 384), (5, 5)])
def test_matches_float64_statistics(shape):
    rng = np.random.default_rng(sum(shape))
    f = (rng.poisson(100, shape) + rng.normal(0, 3, shape)).astype(np.float32)
    m, s = oracle.frame_std(f)
    f64 = f.astype(np.float64)
    assert abs(m - f64.mean()) <= 2e-6 * abs(f64.mean())
    assert abs(s - f64.std(ddof=1)) <= 2e-5 * f64.std(ddof=1)


def test_tree_order_known_answer():
    # 64 pixels: one level of stride 8, then the 8 partial sums in order.
    # x[0] = 2^24 swallows the seven +1 of its group (spacing 2 at 2^24); the
    # other groups sum to 8, which survives: mean = (2^24 + 56) / 64.  A
    # sequential sum would lose every +1: 2^24 / 64.
    x = np.ones(64, np.float32)
    x[0] = 2.0 ** 24
    m, _ = oracle.frame_std(x)
    assert m == np.float32((2.0 ** 24 + 56) / 64)
    assert m != np.float32(2.0 ** 24 / 64)


def test_markers_are_skipped_but_counted():
    rng = np.random.default_rng(3)
    f = rng.normal(50, 4, (40, 40)).astype(np.float32)
    f.reshape(-1)[rng.choice(1600, 37, replace=False)] = MARKER
    m, s = oracle.frame_std(f)
    valid = f.reshape(-1)[~np.isnan(f.reshape(-1))].astype(np.float64)
    N = f.size
    want_m = valid.sum() / N                       # the divisor is the pixel count
    want_s = np.sqrt(((valid - np.float32(m)) ** 2).sum() / (N - 1))
    assert abs(m - want_m) <= 1e-5 * abs(want_m)
    assert abs(s - want_s) <= 1e-4 * want_s


def _tree_np(vals, len0):
    """Independent restatement (vectorised numpy f32) of the in-place levels:
    stride s = ceil(len / 8), arr[i] = sum_j arr[i + s j] (j order), indices
    past len contribute nothing; until len <= 24; then the in-order sum."""
    t = vals.astype(np.float32).copy()
    n = len0
    while n > 24:
        s = (n + 7) // 8
        acc = np.zeros(s, np.float32)
        for j in range(8):
            idx = np.arange(s) + s * j
            ok = idx < n
            add = np.zeros(s, np.float32)
            add[ok] = t[idx[ok]]
            acc = (acc + add).astype(np.float32)
        t[:s] = acc
        n = s
    tot = np.float32(0)
    for i in range(n):
        tot = np.float32(tot + t[i])
    return tot


@pytest.mark.parametrize("n", [25, 200, 1000, 4099, 65536, 147456])
def test_matches_shader_restatement_bit_for_bit(n):
    rng = np.random.default_rng(n)
    x = (rng.poisson(200, n) + rng.normal(0, 2, n)).astype(np.float32)
    x[rng.choice(n, n // 97, replace=False)] = MARKER
    m, s = oracle.frame_std(x)
    xm = np.where(np.isnan(x), np.float32(0), x).astype(np.float32)
    # mean: the out-of-place level (markers skipped), then the in-place levels
    s0 = (n + 7) // 8
    lvl = np.zeros(s0, np.float32)
    for j in range(8):
        idx = np.arange(s0) + s0 * j
        ok = idx < n
        add = np.zeros(s0, np.float32)
        add[ok] = xm[idx[ok]]
        lvl = (lvl + add).astype(np.float32)
    want_m = np.float32(_tree_np(lvl, s0) / np.float32(n))
    assert np.float32(m) == want_m
    d = (xm - want_m).astype(np.float32)
    sq = np.where(np.isnan(x), np.float32(0), (d * d).astype(np.float32)).astype(np.float32)
    var = np.float32(_tree_np(sq, n) / np.float32(n - 1))
    assert np.float32(s) == np.float32(np.sqrt(var))
