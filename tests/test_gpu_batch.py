"""Batched detection of frames of different sizes (one kernel per stage over a
frame table, graph replay, autotuned schedule) against separate one-size runs:
rows and counts must be identical bit for bit, per frame."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


SIZES = [(96, 128), (200, 150), (256, 256), (131, 77), (64, 64)]
PARAMS = dict(separation=7, minmass=400.0, characterize=True)


def _videos(torch, nframes=3):
    from gpu_tracking_amd import synth
    vids = []
    for k, (H, W) in enumerate(SIZES):
        s = synth.spec("full", nrows=H, ncols=W, n_particles=max(4, H * W // 1000))
        v = synth.video_numpy(s, nframes, seed=400 + k)
        vids.append(torch.from_numpy(v.view(np.int16)).cuda())
    return vids


def _separate(torch, vids, cap):
    import gpu_tracking_amd as gt
    res = []
    for v in vids:
        with gt.Context(gt.make_params(9, **PARAMS), v.shape[1], v.shape[2]) as ctx:
            out, counts = ctx.detect_device(v, cap_per_frame=cap)
            torch.cuda.synchronize()
            res.append((out.cpu().numpy(), counts.cpu().numpy()))
    return res


def _frames(vids):
    frames, ids, where = [], [], []
    T = vids[0].shape[0]
    for t in range(T):                        # interleave videos: mixed sizes per launch
        for k, v in enumerate(vids):
            frames.append(v[t])
            ids.append(t)
            where.append((k, t))
    return frames, ids, where


def _check(out, counts, sep, where):
    o = out.cpu().numpy()
    c = counts.cpu().numpy()
    for i, (k, t) in enumerate(where):
        ro, rc = sep[k]
        assert c[i] == rc[t], f"frame {i} (video {k}, t {t}): count {c[i]} vs {rc[t]}"
        # rows arrive in atomic order: compare them ordered by seed (word 10)
        a = o[i, :c[i]].view(np.uint32).reshape(c[i], 12)
        b = ro[t, :rc[t]].view(np.uint32).reshape(rc[t], 12)
        a = a[np.argsort(a[:, 10], kind="stable")]
        b = b[np.argsort(b[:, 10], kind="stable")]
        assert np.array_equal(a, b), f"frame {i} (video {k}, t {t}) rows differ"


@pytest.mark.parametrize("graph,launch_px,slots", [(True, 1 << 28, 256), (False, 1 << 28, 256),
                                                   (True, 100_000, 4)])
def test_batch_matches_separate_runs(torch_dev, graph, launch_px, slots):
    import gpu_tracking_amd as gt
    vids = _videos(torch_dev)
    cap = 2048
    sep = _separate(torch_dev, vids, cap)
    frames, ids, where = _frames(vids)
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=256 * 256,
                  max_launch_pixels=launch_px, max_slots=slots) as b:
        b.set_graph(graph)
        out, counts = b.detect(frames, frame_ids=ids, cap_per_frame=cap)
        torch_dev.cuda.synchronize()
        _check(out, counts, sep, where)
        if launch_px < (1 << 28):
            assert b.schedule()["launches_last_call"] > 1


@pytest.mark.parametrize("lanes,graph", [(2, True), (3, True), (4, False)])
def test_batch_lanes_match_separate_runs(torch_dev, lanes, graph):
    # launches dealt to several streams; repeated calls replay each lane's graph
    import gpu_tracking_amd as gt
    vids = _videos(torch_dev)
    cap = 2048
    sep = _separate(torch_dev, vids, cap)
    frames, ids, where = _frames(vids)
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=256 * 256,
                  max_launch_pixels=100_000, max_slots=4) as b:
        b.set_graph(graph)
        b.set_lanes(lanes)
        for rep in range(2):
            out, counts = b.detect(frames, frame_ids=ids, cap_per_frame=cap)
            torch_dev.cuda.synchronize()
            _check(out, counts, sep, where)
        sch = b.schedule()
        assert sch["lanes"] == lanes and sch["launches_last_call"] > lanes


def test_batch_graph_replays_stay_exact(torch_dev):
    # several calls (each capturing and replaying a graph) with eager work in
    # between; every replay must equal the separate runs
    import gpu_tracking_amd as gt
    vids = _videos(torch_dev, nframes=2)
    cap = 2048
    sep = _separate(torch_dev, vids, cap)
    frames, ids, where = _frames(vids)
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=256 * 256,
                  max_launch_pixels=150_000, max_slots=8) as b:
        for rep in range(3):
            out, counts = b.detect(frames, frame_ids=ids, cap_per_frame=cap)
            torch_dev.cuda.synchronize()
            _check(out, counts, sep, where)
            junk = torch_dev.ones(1 << 20, device="cuda") * rep    # eager work between replays
            torch_dev.cuda.synchronize()
            del junk


def test_batch_autotune_keeps_results(torch_dev):
    import gpu_tracking_amd as gt
    vids = _videos(torch_dev, nframes=2)
    cap = 2048
    sep = _separate(torch_dev, vids, cap)
    frames, ids, where = _frames(vids)
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=256 * 256) as b:
        ms = b.tune(frames, cap_per_frame=cap)
        assert ms > 0
        sch = b.schedule()
        assert sch["pixels_per_launch"] >= 256 * 256 and sch["walk_blocks"] >= 1
        out, counts = b.detect(frames, frame_ids=ids, cap_per_frame=cap)
        torch_dev.cuda.synchronize()
        _check(out, counts, sep, where)


def test_batch_rejects_bad_input(torch_dev):
    import gpu_tracking_amd as gt
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=64 * 64) as b:
        big = torch_dev.zeros((128, 128), dtype=torch_dev.int16, device="cuda")
        with pytest.raises(gt.GtError):
            b.detect([big])
        f32 = torch_dev.zeros((32, 32), dtype=torch_dev.float32, device="cuda")
        with pytest.raises(TypeError):
            b.detect([f32])


# --------------------------------------------------------------------------
# The batched workload's own frame sizes in one call, every frame checked
# against the oracle (not only against one-size HIP runs).  Sizes (rows,
# columns) and generator as the batched workload: the full workload's
# generator at 1 particle per 1000 px^2, its detection parameters.
WORKLOAD_SIZES = [(256, 256), (512, 384), (1024, 1024), (2048, 1024)]
FIELDS = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]


@pytest.mark.parametrize("graph,launch_px", [(True, 3 << 20), (False, 3 << 20), (True, 16 << 20)])
def test_batch_workload_sizes_match_the_oracle(torch_dev, oracle, graph, launch_px):
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    frames, ids, host = [], [], []
    for k, (H, W) in enumerate(WORKLOAD_SIZES):
        s = synth.spec("full", nrows=H, ncols=W, n_particles=H * W // 1000)
        v = synth.video_numpy(s, 2, seed=20261203 + 4000 + k)
        d = torch_dev.from_numpy(v.view(np.int16)).cuda()
        for t in range(2):
            frames.append(d[t])
            ids.append(t)
            host.append(v[t])
    # interleave sizes so that launches mix them
    order = [0, 2, 4, 6, 1, 3, 5, 7]
    frames = [frames[i] for i in order]
    ids = [ids[i] for i in order]
    host = [host[i] for i in order]
    print(f"[batch oracle] graph={graph} sizes {[f.shape for f in frames]}", flush=True)
    cap = 8192
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=2048 * 1024,
                  max_launch_pixels=launch_px, max_slots=8) as b:
        b.set_graph(graph)
        out, counts = b.detect(frames, frame_ids=ids, cap_per_frame=cap)
        torch_dev.cuda.synchronize()
        # 6.6 Mpx in all: several launches under a 3 Mpx budget, one launch
        # per stage (one tile queue over all four sizes) under 16 Mpx
        launches = b.schedule()["launches_last_call"]
        assert launches > 1 if launch_px < (6 << 20) else launches == 1
        got = gt.detections_to_numpy(out, counts)
    for i, fr in enumerate(host):
        H, W = fr.shape
        p = oracle.derive_trackpy(H, W, diameter=9, separation=7, minmass=400.0)
        ref, _ = oracle.detect_frame(fr, p, characterize=True)
        ref = np.sort(ref, order="seed")
        g = got[i]
        assert len(g) == len(ref) and len(ref) > 0, (i, fr.shape, len(g), len(ref))
        assert np.array_equal(g["seed"], ref["seed"]), i
        assert np.all(g["frame"] == ids[i])
        for f in FIELDS:
            assert np.array_equal(g[f].view(np.uint32), ref[f].view(np.uint32)), (i, f)


def test_batch_many_sizes_split_by_size_runs(torch_dev):
    # ten frame sizes in one call: a launch holds at most 8 runs of equal
    # size, so the call takes two launches even though its pixels fit one;
    # rows equal separate one-size runs
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    sizes = [(64 + 16 * k, 96 + 8 * k) for k in range(10)]
    vids = []
    for k, (H, W) in enumerate(sizes):
        sp = synth.spec("full", nrows=H, ncols=W, n_particles=max(4, H * W // 1000))
        vids.append(torch_dev.from_numpy(synth.video_numpy(sp, 1, seed=900 + k).view(np.int16)).cuda())
    cap = 2048
    sep = _separate(torch_dev, vids, cap)
    frames = [v[0] for v in vids]
    where = [(k, 0) for k in range(len(vids))]
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=1 << 20,
                  max_launch_pixels=64 << 20, max_slots=64) as b:
        out, counts = b.detect(frames, frame_ids=[0] * len(frames), cap_per_frame=cap)
        torch_dev.cuda.synchronize()
        sch = b.schedule()
        assert sch["launches_last_call"] == 2
        # the first launch holds 8 runs of one frame each; the grid covers it
        assert sch["max_size_runs_per_launch_last_call"] == 8
        assert sch["frames_per_launch_last_call"] == 8
        assert sch["tile_shape"] == (32, 64)
        # a launch of 8 frames has at least the 8 smallest frames' tiles
        tiles = sorted(((H + 31) // 32) * ((W + 63) // 64) for H, W in sizes)
        assert sch["tile_blocks_last_call"] >= sum(tiles[:8])
        assert sch["tile_blocks_last_call"] % 8 == 0
    _check(out, counts, sep, where)


def test_batch_videos_as_3d_tensors(torch_dev):
    # 3-D videos in the list: their frames join the call in order, with the
    # rows of separate one-size runs; a 2-D frame may sit between them
    import gpu_tracking_amd as gt
    vids = _videos(torch_dev)
    cap = 2048
    sep = _separate(torch_dev, vids, cap)
    items = [vids[0], vids[1][1], vids[2], vids[3], vids[4]]
    where = ([(0, t) for t in range(3)] + [(1, 1)] + [(2, t) for t in range(3)] +
             [(3, t) for t in range(3)] + [(4, t) for t in range(3)])
    ids = [t for _, t in where]
    print(f"[batch 3-D] {len(where)} frames from {len(items)} items", flush=True)
    with gt.Batch(gt.make_params(9, **PARAMS), max_frame_pixels=256 * 256,
                  max_launch_pixels=1 << 20, max_slots=64) as b:
        out, counts = b.detect(items, frame_ids=ids, cap_per_frame=cap)
        torch_dev.cuda.synchronize()
        assert out.shape[0] == len(where)
    _check(out, counts, sep, where)
