"""The oracle's adaptive_background detection (or_detect_frame_adaptive)
pinned on the CPU: one pass equals the plain detection (rough factor 0.5)
filtered by the adaptive rule; two passes equal an independent numpy
restatement of the masking and the re-estimated noise; passes only add
rows."""
import numpy as np
import pytest

import oracle

MARKER = np.array([0xFFFFFFFF], np.uint32).view(np.float32)[0]


def _frame(seed, H=160, W=176, n=40):
    from gpu_tracking_amd import synth
    s = synth.spec("full", nrows=H, ncols=W, n_particles=n)
    return synth.video_numpy(s, 1, seed)[0].astype(np.float32)


def _params(H, W, **kw):
    return oracle.derive_trackpy(H, W, diameter=9, separation=7, adaptive_background=True, **kw)


def _rule(rows, p, std):
    mm = p.minmass if p.minmass > 0 else np.float32(p.minmass_snr) * rows["count"] * np.float32(std)
    return ~((rows["mass"] < mm) | (rows["max_intensity"] < np.float32(p.snr) * np.float32(std)))


def _mask(fr, p, rows):
    rint = (p.circle - 1) // 2
    r2 = np.float32(rint) * np.float32(rint)
    H, W = fr.shape
    for row in rows:
        u, v = int(np.rint(row["x"])), int(np.rint(row["y"]))
        for i in range(-rint, rint + 1):
            for j in range(-rint, rint + 1):
                if np.float32(i * i) + np.float32(j * j) > r2:
                    continue
                if 0 <= u + i < H and 0 <= v + j < W:
                    fr[u + i, v + j] = MARKER


def test_std_div_reduces_to_frame_std():
    f = _frame(1)
    assert oracle.frame_std_div(f, f.size, f.size - 1) == oracle.frame_std(f)


@pytest.mark.parametrize("kw", [dict(snr=1.0, minmass_snr=0.2), dict(minmass=300.0, snr=2.0)])
def test_one_pass_is_the_filtered_plain_detection(kw):
    f = _frame(2)
    H, W = f.shape
    p = _params(H, W, **kw)
    got = np.sort(oracle.detect_frame_adaptive(f, p, 1), order="seed")
    _, std = oracle.frame_std(f)
    p.image_std = std
    rows, _ = oracle.detect_frame(f, p, characterize=True)
    want = np.sort(rows[_rule(rows, p, std)], order="seed")
    assert len(want) > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_two_passes_match_a_restatement():
    f = _frame(3, n=80)
    H, W = f.shape
    p = _params(H, W, snr=1.5, minmass_snr=0.3)
    _, std0 = oracle.frame_std(f)
    p.image_std = std0
    rows, _ = oracle.detect_frame(f, p, characterize=True)
    ok1 = _rule(rows, p, std0)
    acc1, rest = rows[ok1], rows[~ok1]
    masked = f.copy()
    _mask(masked, p, acc1)
    div = f.size - int(acc1["count"].astype(np.uint32).sum())
    # the marked pixels are the accepted disks (union), in the frame
    assert np.isnan(masked).sum() > 0
    _, std1 = oracle.frame_std_div(masked, div, div - 1)
    ok2 = _rule(rest, p, std1)
    want = np.sort(np.concatenate([acc1, rest[ok2]]), order="seed")
    got = np.sort(oracle.detect_frame_adaptive(f, p, 2), order="seed")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert std1 < std0       # the brightest pixels left the estimate


def test_passes_only_add_rows():
    f = _frame(4, n=90)
    H, W = f.shape
    p = _params(H, W, snr=2.0, minmass_snr=0.5)
    prev = set()
    for n in range(4):
        s = set(oracle.detect_frame_adaptive(f, p, n)["seed"].tolist())
        assert prev <= s
        prev = s
    assert len(prev) > 0


# ---- LoG style: rows carry r > 0 and mark disks of radius ceil(r) -------------

def _log_case(seed, n=60):
    from gpu_tracking_amd import synth
    s = synth.spec("log", nrows=128, ncols=144, n_particles=n)
    f = synth.video_numpy(s, 1, seed)[0].astype(np.float32)
    p = oracle.derive_log(128, 144, 4.0)
    p.snr, p.minmass_snr, p.rough_snr_factor = 1.5, 0.3, 0.5
    return f, p, oracle.log_radii(1.5, 4.0, 5, False)


def _log_rule(rows, p, std):
    mm = np.float32(p.minmass_snr) * rows["count"] * np.float32(std)
    return ~((rows["mass"] < mm) | (rows["max_intensity"] < np.float32(p.snr) * np.float32(std)))


def _mask_r(fr, rows):
    H, W = fr.shape
    for row in rows:
        rint = int(np.ceil(row["r"]))
        r2 = np.float32(rint) * np.float32(rint)
        u, v = int(np.rint(row["x"])), int(np.rint(row["y"]))
        for i in range(-rint, rint + 1):
            for j in range(-rint, rint + 1):
                if np.float32(i * i) + np.float32(j * j) > r2:
                    continue
                if 0 <= u + i < H and 0 <= v + j < W:
                    fr[u + i, v + j] = MARKER


def test_log_one_pass_is_the_filtered_plain_detection():
    f, p, radii = _log_case(5)
    got = np.sort(oracle.log_detect_frame_adaptive(f, p, radii, 1, overlap_threshold=1.0),
                  order="seed")
    _, std = oracle.frame_std(f)
    p.image_std = std
    rows, _ = oracle.log_detect_frame(f, p, radii, characterize=True, overlap_threshold=1.0)
    want = np.sort(rows[_log_rule(rows, p, std)], order="seed")
    assert len(want) > 0 and np.all(want["r"] > 0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_log_two_passes_match_a_restatement():
    f, p, radii = _log_case(6, n=90)
    _, std0 = oracle.frame_std(f)
    p.image_std = std0
    rows, _ = oracle.log_detect_frame(f, p, radii, characterize=True, overlap_threshold=1.0)
    ok1 = _log_rule(rows, p, std0)
    acc1, rest = rows[ok1], rows[~ok1]
    masked = f.copy()
    _mask_r(masked, acc1)
    div = f.size - int(acc1["count"].astype(np.uint32).sum())
    _, std1 = oracle.frame_std_div(masked, div, div - 1)
    ok2 = _log_rule(rest, p, std1)
    want = np.sort(np.concatenate([acc1, rest[ok2]]), order="seed")
    got = np.sort(oracle.log_detect_frame_adaptive(f, p, radii, 2, overlap_threshold=1.0),
                  order="seed")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # larger blobs mark wider disks than the fixed walk radius would
    assert np.any(np.ceil(acc1["r"]) >= 3)


def test_log_passes_only_add_rows():
    f, p, radii = _log_case(7, n=90)
    prev = set()
    for n in range(4):
        s = set(oracle.log_detect_frame_adaptive(f, p, radii, n, overlap_threshold=1.0)["seed"]
                .tolist())
        assert prev <= s
        prev = s
    assert len(prev) > 0
