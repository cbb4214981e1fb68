"""GPU parity of the LoG style: the HIP path (k_log_tile + k_walk_log through
the C ABI) against the CPU oracle on the same seeded frames.  Candidate
seeds (scale * H * W + pixel) and every row field must be bit-identical;
the end-to-end LoG() DataFrame must equal the oracle's pruned rows."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


def _video(T, seed, H=192, W=224, n=40, noise="gauss:5", **kw):
    from gpu_tracking_amd import synth
    s = synth.spec("log", nrows=H, ncols=W, n_particles=n, noise=noise, **kw)
    return synth.video_numpy(s, T, seed)


def _gpu_rows(torch, video, min_r, max_r, cap=None, **kw):
    import gpu_tracking_amd as gt
    T, H, W = video.shape
    p = gt.make_log_params(min_r, max_r, **kw)
    with gt.Context(p, H, W) as ctx:
        t = torch.from_numpy(video.view(np.int16) if video.dtype == np.uint16 else video).cuda()
        out, counts = ctx.detect_device(t, cap_per_frame=cap or H * W)
        torch.cuda.synchronize()
        return gt.detections_to_numpy(out, counts)


def _oracle_rows(oracle, frame, min_r, max_r, n_radii=10, log_spacing=False, characterize=True,
                 overlap_threshold=1.0, **kw):
    H, W = frame.shape
    p = oracle.derive_log(H, W, max_r, **kw)
    radii = oracle.log_radii(min_r, max_r, n_radii, log_spacing)
    rows, nc = oracle.log_detect_frame(frame, p, radii, characterize=characterize,
                                       overlap_threshold=overlap_threshold)
    return rows, nc


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
    # (frames, seed, frame kwargs, LoG kwargs)
    (2, 1, dict(), dict(min_r=1.5, max_r=6.0, n_radii=10, log_spacing=True)),
    (2, 2, dict(H=160, W=136, n=25), dict(min_r=1.5, max_r=6.0, n_radii=10, log_spacing=False)),
    (1, 3, dict(H=97, W=131, n=12), dict(min_r=2.0, max_r=4.0, n_radii=3, log_spacing=False)),
    (1, 4, dict(H=128, W=128, n=15), dict(min_r=2.5, max_r=2.5, n_radii=1)),
    (1, 5, dict(H=150, W=150, n=15), dict(min_r=1.0, max_r=9.0, n_radii=6, log_spacing=True,
                                          minmass=300.0, max_iterations=4)),
]


@pytest.mark.parametrize("T,seed,fkw,lkw", CASES)
def test_log_rows_bit_identical(torch_dev, oracle, T, seed, fkw, lkw):
    print(f"[log case] T={T} seed={seed} {fkw} {lkw}", flush=True)
    video = _video(T, seed, **fkw)
    lkw = dict(lkw)
    min_r, max_r = lkw.pop("min_r"), lkw.pop("max_r")
    got = _gpu_rows(torch_dev, video, min_r, max_r, **lkw)
    okw = {k: v for k, v in lkw.items() if k in ("n_radii", "log_spacing")}
    pkw = {k: v for k, v in lkw.items() if k in ("minmass", "max_iterations")}
    for t in range(T):
        ref, nc = _oracle_rows(oracle, video[t], min_r, max_r, **okw, **pkw)
        assert nc > 0
        _assert_rows_equal(got[t], ref, f"frame {t}")


def test_log_uint16_frames(torch_dev, oracle):
    video = _video(1, 6, H=128, W=144, n=20, noise="poisson", background=(100.0, 120.0))
    assert video.dtype == np.uint16
    got = _gpu_rows(torch_dev, video, 1.5, 5.0, n_radii=5)
    ref, _ = _oracle_rows(oracle, video[0].astype(np.float32), 1.5, 5.0, n_radii=5)
    _assert_rows_equal(got[0], ref, "u16")


def test_log_end_to_end_dataframe(torch_dev, oracle):
    import gpu_tracking_amd as gt
    video = _video(2, 7, H=160, W=176, n=30)
    df = gt.LoG(video, 1.5, 6.0, n_radii=10, log_spacing=True)
    assert list(df.columns) == ["frame", "y", "x", "mass", "r", "Rg", "raw", "signal", "ecc",
                                "raw_mass", "raw_bg_median", "raw_mass_corrected"]
    for t in range(2):
        ref, _ = _oracle_rows(oracle, video[t], 1.5, 6.0, n_radii=10, log_spacing=True,
                              overlap_threshold=0.0)
        sub = df[df.frame == t]
        assert len(sub) == len(ref)
        for col, f in [("y", "x"), ("x", "y"), ("mass", "mass"), ("r", "r"), ("Rg", "Rg"),
                       ("raw", "raw_mass"), ("signal", "signal"), ("ecc", "ecc")]:
            a = sub[col].to_numpy(np.float32)
            b = ref[f]
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) or \
                np.array_equal(a, b, equal_nan=True), col
        # doughnut per detection at its own radius, ring [r, 2r]
        for i in range(min(len(ref), 50)):
            d = oracle.doughnut(video[t], ref[i:i + 1], float(ref["r"][i]),
                                2.0 * float(ref["r"][i]), 0.0)[0]
            got = sub[["raw_mass", "raw_bg_median", "raw_mass_corrected"]].to_numpy(np.float32)[i]
            assert np.array_equal(got, d), i


def test_log_overlap_threshold_variants(torch_dev, oracle):
    import gpu_tracking_amd as gt
    video = _video(1, 8, H=128, W=128, n=25)
    for thr in (0.5, 1.0):
        df = gt.LoG(video, 1.5, 4.0, n_radii=5, overlap_threshold=thr, doughnut_correction=False,
                    characterize=False)
        assert list(df.columns) == ["frame", "y", "x", "mass", "r"]
        ref, _ = _oracle_rows(oracle, video[0], 1.5, 4.0, n_radii=5, overlap_threshold=thr)
        assert len(df) == len(ref)
        assert np.array_equal(df["mass"].to_numpy(np.float32), ref["mass"])


def test_log_tile_buffer_spill_bit_identical(torch_dev, oracle):
    # pure noise at 20 small scales: about 1360 scale-space maxima per
    # 128 x 16 tile, more than a tile buffers, so candidates spill to the top
    # of the list and are walked from global memory; the rest walk from the
    # tiles' staged planes.  Rows equal the oracle's either way.
    rng = np.random.default_rng(5)
    frame = (50.0 + rng.normal(0.0, 5.0, (64, 256))).astype(np.float32)
    print("[log spill] 64x256 noise, 20 scales", flush=True)
    ref, nc = _oracle_rows(oracle, frame, 1.0, 3.0, n_radii=20)
    assert nc > 8 * 1024
    got = _gpu_rows(torch_dev, frame[None], 1.0, 3.0, n_radii=20, log_spacing=False)
    _assert_rows_equal(got[0], ref, "spill")


def test_log_capacity_overflow_reported(torch_dev):
    video = _video(1, 9, H=128, W=128, n=20)
    import gpu_tracking_amd as gt
    p = gt.make_log_params(1.5, 6.0)
    with gt.Context(p, 128, 128) as ctx:
        out, counts = ctx.detect_device(torch_dev.from_numpy(video).cuda(), cap_per_frame=8)
        torch_dev.cuda.synchronize()
        assert (int(counts[0].item()) & 0xFFFFFFFF) > 8


# --------------------------------------------------------------------------
# The LoG workload at its own shape: one 2048 x 2048 f32 frame of the LoG
# workload generator (3000 particles, sigma U[1, 5], amplitude U[200, 2000],
# background 50 + N(0, 5^2)); radii 1.5 .. 6, 10 log-spaced scales; without
# and with threshold_rel = 0.05 (no reference counterpart).
@pytest.fixture(scope="module")
def log_workload(oracle):
    from concurrent.futures import ThreadPoolExecutor
    from gpu_tracking_amd import synth
    frame = synth.video_numpy(synth.spec("log"), 1, 31)
    radii = oracle.log_radii(1.5, 6.0, 10, True)

    def ref(rel):
        p = oracle.derive_log(2048, 2048, 6.0, threshold_rel=rel)
        return oracle.log_detect_frame(frame[0], p, radii, overlap_threshold=1.0)

    # the oracle's C code releases the GIL: both variants at once
    with ThreadPoolExecutor(2) as ex:
        refs = dict(zip((0.0, 0.05), ex.map(ref, (0.0, 0.05))))
    return frame, refs


@pytest.mark.parametrize("rel", [0.0, 0.05])
def test_log_workload_2048_bit_identical(torch_dev, log_workload, rel):
    print(f"[log workload] 2048x2048 f32, threshold_rel={rel}", flush=True)
    frame, refs = log_workload
    ref, nc = refs[rel]
    got = _gpu_rows(torch_dev, frame, 1.5, 6.0, n_radii=10, log_spacing=True, threshold_rel=rel)
    _assert_rows_equal(got[0], ref, f"threshold_rel={rel}")
    if rel > 0:
        # the rule keeps the particle-sized maxima and drops the noise
        assert 1000 < len(ref) < len(refs[0.0][0]) // 10
