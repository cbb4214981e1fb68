"""The reference's Python entry points (batch, characterize_points) end to end
through gt_execute: host staging, GPU detection, host post-processing
(filter_close, doughnut photometry) and column assembly, against the oracle
run frame by frame on the same video."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gt(gpu_ok):
    import gpu_tracking_amd
    return gpu_tracking_amd


def _video(T=4, seed=3, H=200, W=232, n=30):
    from gpu_tracking_amd import synth
    return synth.video_numpy(synth.spec("small", nrows=H, ncols=W, n_particles=n), T, seed)


def _expected(oracle, video, diameter, separation, minmass, characterize, doughnut,
              filter_close=True, frames=None, bg_radius=None, gap_radius=0.0):
    T, H, W = video.shape
    p = oracle.derive_trackpy(H, W, diameter=diameter, separation=separation, minmass=minmass)
    out = []
    for t in (range(T) if frames is None else frames):
        rows, _ = oracle.detect_frame(video[t], p, characterize=characterize)
        rows = np.sort(rows, order="seed")
        if filter_close:
            rows = rows[oracle.filter_close(rows, float(separation))]
        cols = [np.full(len(rows), t, np.float32), rows["x"], rows["y"], rows["mass"]]
        if characterize:
            cols += [rows["Rg"], rows["raw_mass"], rows["signal"], rows["ecc"]]
        if doughnut:
            d = oracle.doughnut(video[t], rows, diameter / 2.0,
                                float(diameter) if bg_radius is None else bg_radius, gap_radius)
            cols += [d[:, 0], d[:, 1], d[:, 2]]
        out.append(np.stack(cols, 1).astype(np.float32) if len(rows) else
                   np.zeros((0, len(cols)), np.float32))
    return np.concatenate(out, 0)


def _frame_matrix(df):
    return df.to_numpy(dtype=np.float64).astype(np.float32)


def test_batch_default_columns_and_values(gt, oracle):
    v = _video()
    df = gt.batch(v, 9, minmass=300)
    assert list(df.columns) == ["frame", "y", "x", "mass", "Rg", "raw", "signal", "ecc",
                                "raw_mass", "raw_bg_median", "raw_mass_corrected"]
    assert str(df["frame"].dtype) == "int64" and str(df["y"].dtype) == "float32"
    ref = _expected(oracle, v, 9, 10, 300.0, True, True)
    got = _frame_matrix(df)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_batch_without_characterize_or_doughnut(gt, oracle):
    v = _video(T=3, seed=8)
    df = gt.batch(v, 7, separation=6, minmass=150, characterize=False,
                  doughnut_correction=False, filter_close=False)
    assert list(df.columns) == ["frame", "y", "x", "mass"]
    ref = _expected(oracle, v, 7, 6, 150.0, False, False, filter_close=False)
    assert np.array_equal(_frame_matrix(df).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("kw,dough,fc,dtype", [
    (dict(doughnut_correction=False), False, True, np.uint16),
    (dict(filter_close=False), True, False, np.uint16),
    (dict(bg_radius=7.0, gap_radius=1.5), True, True, np.uint16),
    (dict(), True, True, np.float32),
    (dict(), True, True, np.uint8),
])
def test_batch_post_processing_variants(gt, oracle, kw, dough, fc, dtype):
    # filter_close and the doughnut run on the GPU; both must equal the
    # oracle's host restatement bit for bit
    v = _video(T=3, seed=12, n=60)
    if dtype == np.uint8:
        v = (v // 64).astype(np.uint8)
    else:
        v = v.astype(dtype)
    mm = 20.0 if dtype == np.uint8 else 300.0
    df = gt.batch(v, 9, minmass=mm, **kw)
    ref = _expected(oracle, v, 9, 10, mm, True, dough, filter_close=fc,
                    bg_radius=kw.get("bg_radius"), gap_radius=kw.get("gap_radius", 0.0))
    got = _frame_matrix(df)
    assert got.shape == ref.shape and len(got) > 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_batch_frames_subset_and_order(gt, oracle):
    v = _video(T=5, seed=9)
    df = gt.batch(v, 9, minmass=300, frames=[4, 1, 2])
    assert df["frame"].drop_duplicates().tolist() == [4, 1, 2]
    ref = _expected(oracle, v, 9, 10, 300.0, True, True, frames=[4, 1, 2])
    assert np.array_equal(_frame_matrix(df).view(np.uint32), ref.view(np.uint32))


def test_batch_errors(gt):
    v = _video(T=2)
    with pytest.raises(gt.GtError) as e:
        gt.batch(v, 9, frames=[0, 7])
    assert e.value.status == 7  # FrameOutOfBounds
    # the message of the call's own thread, not "no error"
    assert "outside the length" in str(e.value)
    with pytest.raises(gt.GtError) as e:
        gt.batch(v[:0], 9)
    assert e.value.status == 5  # EmptyIterator


def test_batch_accepts_2d_and_float_input(gt, oracle):
    v = _video(T=1, seed=12)
    a = gt.batch(v[0], 9, minmass=300)
    b = gt.batch(v.astype(np.float64), 9, minmass=300)   # cast to f32 like the reference
    assert np.array_equal(_frame_matrix(a).view(np.uint32), _frame_matrix(b).view(np.uint32))


def test_characterize_points_matches_oracle(gt, oracle):
    import pandas as pd
    v = _video(T=3, seed=14)
    df = gt.batch(v, 9, minmass=300, doughnut_correction=False)
    pts = df[["frame", "y", "x"]].copy()
    pts["y"] += 0.3
    out = gt.characterize_points(v, pts, diameter=9, doughnut_correction=False)
    assert list(out.columns) == ["frame", "y", "x", "mass", "Rg", "raw", "signal", "ecc"]
    assert len(out) == len(pts)
    T, H, W = v.shape
    p = oracle.derive_trackpy(H, W, diameter=9)
    exp = []
    for t in range(T):
        sel = pts[pts["frame"] == t]
        rows = np.zeros(len(sel), oracle.ROW_DTYPE)
        rows["x"] = sel["y"].to_numpy(np.float32)
        rows["y"] = sel["x"].to_numpy(np.float32)
        P = oracle.bandpass(v[t], p)
        rows = oracle.characterize(P, v[t], p, rows, use_f64=True, recompute_mass=True)
        exp.append(np.stack([np.full(len(rows), t, np.float32), rows["x"], rows["y"], rows["mass"],
                             rows["Rg"], rows["raw_mass"], rows["signal"], rows["ecc"]], 1))
    exp = np.concatenate(exp).astype(np.float32)
    got = _frame_matrix(out)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_characterize_points_without_frames_and_with_r(gt, oracle):
    v = _video(T=2, seed=15)
    pts = np.array([[50.2, 60.7, 3.0], [100.5, 120.5, 4.5], [20.0, 30.0, 2.2]], np.float32)
    out = gt.characterize_points(v, pts, has_frames=False, has_r=True,
                                 doughnut_correction=False)
    # every frame gets every point; diameter defaults to 2*int(max r + 0.5) + 1
    assert list(out.columns) == ["frame", "y", "x", "mass", "r", "Rg", "raw", "signal", "ecc"]
    assert out["frame"].tolist() == [0, 0, 0, 1, 1, 1]
    assert np.allclose(out["r"].to_numpy()[:3], pts[:, 2])
    p = oracle.derive_trackpy(v.shape[1], v.shape[2], diameter=2 * int(4.5 + 0.5) + 1)
    rows = np.zeros(3, oracle.ROW_DTYPE)
    rows["x"], rows["y"], rows["r"] = pts[:, 0], pts[:, 1], pts[:, 2]
    P = oracle.bandpass(v[1], p)
    ref = oracle.characterize(P, v[1], p, rows, use_f64=True, recompute_mass=True)
    got = out[out["frame"] == 1]
    assert np.array_equal(got["mass"].to_numpy(np.float32), ref["mass"])
    assert np.array_equal(got["ecc"].to_numpy(np.float32), ref["ecc"])


def test_characterize_points_unsorted_frames_rejected(gt):
    v = _video(T=3, seed=16)
    pts = np.array([[1, 50, 50], [0, 60, 60]], np.float32)
    with pytest.raises(gt.GtError) as e:
        gt.characterize_points(v, pts, diameter=9, has_frames=True)
    assert e.value.status == 6  # NonSortedCharacterization


def test_batch_with_linking(gt, oracle):
    # search_range adds particle and duration columns (column_names,
    # execute_gpu.rs:482-519; durations :1165-1181); ids must equal the
    # oracle linker run on the oracle's per-frame rows in the same order
    from oracle import linker as OL
    v = _video(T=8, seed=12, n=40)
    df = gt.batch(v, 9, minmass=300, search_range=5.0, memory=2, doughnut_correction=False)
    assert list(df.columns)[-2:] == ["particle", "duration"]
    ref = _expected(oracle, v, 9, 10, 300.0, True, False)
    got = _frame_matrix(df)
    assert np.array_equal(got[:, :-2].view(np.uint32), ref.view(np.uint32))
    ids = OL.link_frames(ref[:, :3], 5.0, 2)
    assert df["particle"].tolist() == ids
    # duration: frames from first to last appearance of the particle, inclusive
    first, last = {}, {}
    for f, i in zip(df["frame"].tolist(), ids):
        first.setdefault(i, f)
        last[i] = f
    assert df["duration"].tolist() == [last[i] - first[i] + 1 for i in ids]


def test_batch_read_back_beyond_the_first_copy(gt, oracle):
    # gt_execute reads a chunk's packed rows back with its kernels, as many as
    # recent chunks held (65536 at first), and the rest in a second copy: a
    # 64-frame chunk of noise maxima (thousands per frame) takes that path,
    # the 6-frame chunk after it the grown first copy
    v = _video(T=70, seed=17, H=512, W=512, n=200)
    print(f"[readback] batch() over {v.shape}", flush=True)
    df = gt.batch(v, 7, separation=6, minmass=0, characterize=False,
                  doughnut_correction=False, filter_close=False)
    ref = _expected(oracle, v, 7, 6, 0.0, False, False, filter_close=False)
    got = _frame_matrix(df)
    assert len(got) > 65536 + 64 and (got[:, 0] < 64).sum() > 65536
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_batch_linking_too_dense_raises(gt):
    # a search range that joins every particle of a crowded frame into one
    # subnetwork: TooDenseToLink (linking.rs: load above 5000), raised from
    # the linking thread at the end of the run, with the reference's message
    from gpu_tracking_amd import _native as N
    v = _video(T=3, seed=14, H=128, W=128, n=90)
    print("[dense] batch() over", v.shape, flush=True)
    with pytest.raises(N.GtError) as e:
        gt.batch(v, 9, minmass=0, search_range=200.0, doughnut_correction=False)
    assert e.value.status == 18
    assert "computational load during linking is abnormally high" in str(e.value)
    # the same context still links a sparse video afterwards
    df = gt.batch(_video(T=3, seed=15, n=20), 9, minmass=300, search_range=5.0,
                  doughnut_correction=False)
    assert list(df.columns)[-2:] == ["particle", "duration"]


def test_batch_linker_reset_points(gt, oracle):
    v = _video(T=6, seed=13, n=25)
    df = gt.batch(v, 9, minmass=300, search_range=5.0, linker_reset_points=[3],
                  doughnut_correction=False, characterize=False)
    a = df[df.frame < 3]
    b = df[df.frame >= 3]
    # ids start over at the reset point
    assert b["particle"].min() == 0 and a["particle"].min() == 0
    assert set(b[b.frame == 3]["particle"]) == set(range(len(b[b.frame == 3])))


def test_batch_across_staging_slots_matches_the_oracle(gt, oracle):
    # gt_execute stages host frames 64 at a time in two alternating pinned
    # slots; 150 frames take three slots, so slot reuse, the GPU doughnut and
    # filter_close buffers and the linker's state all cross slot boundaries
    from oracle import linker as OL
    v = _video(T=150, seed=21, H=128, W=128, n=14)
    print(f"[slots] batch() over {v.shape}", flush=True)
    df = gt.batch(v, 9, minmass=300, search_range=5.0, memory=2)
    assert list(df.columns) == ["frame", "y", "x", "mass", "Rg", "raw", "signal", "ecc",
                                "raw_mass", "raw_bg_median", "raw_mass_corrected", "particle",
                                "duration"]
    ref = _expected(oracle, v, 9, 10, 300.0, True, True)
    got = _frame_matrix(df)
    assert got.shape[0] == ref.shape[0] and got.shape[0] > 150
    assert set(df["frame"]) >= {0, 63, 64, 127, 128, 149}
    assert np.array_equal(got[:, :-2].view(np.uint32), ref.view(np.uint32))
    assert df["particle"].tolist() == OL.link_frames(ref[:, :3], 5.0, 2)
