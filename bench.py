"""Throughput of the detection hot path on one MI355X.

Workload (default "full"): 2000 synthetic 1024x1024 uint16 frames, 1000
particles per frame (sigma ~ U[1.2, 2], A ~ lognormal(ln 600, 0.5), Brownian
D = 1, 5% blinking, background ramp 80..140 + Poisson), parameters
diameter 9, separation 7, minmass 400, characterize on.  The frames are
rendered on the GPU and stay in HBM; one step is one pass of the pipeline
(bandpass -> local max -> walk + characterize) over all frames.

Prints ONE JSON line: frames/s, the roofline of the dominant kernel (from
HIP events around each stage on the launch stream), and the CPU oracle's
frames/s on a bounded sample of the same frames on the host cores, plus a
bit-for-bit parity check of that sample.

    python bench.py [--steps K] [--warmup W] [--workload full|small|log|batch]

Workload "batch": 64 seeded videos, frame size drawn from {256x256, 512x384,
1024x1024, 2048x1024}, length uniform in [200, 1000] frames (scaled by
--batch-scale), u16, full-workload generator at 1 particle per 1000 px^2, the
full workload's parameters; all videos go through one batch context
(persistent kernels over a frame table, graph replay, autotuned grids) and,
for comparison, through one one-size context per video; the two must agree
per frame (counts and an order-independent checksum of every row).

Workload "learned": the full workload's frames and parameters, detected with
the learned centroid proposal (a bf16 CNN on the matrix cores proposes where
each candidate's walk ends; one walk evaluation verifies it; the full walk
runs where it fails).  The network is first trained on the GPU on about 2M
(patch, walked end position) pairs from frames of the same generator with
disjoint seeds.  Reports learned and plain frames/s, the fraction of
particles within 1e-3 px of the full walk, and the network kernel's MFMA
roofline.

Workload "log": 500 synthetic 2048x2048 float32 frames, 3000 particles per
frame (sigma ~ U[1, 5], A ~ U[200, 2000], background 50 + N(0, 5^2)), LoG
style with radii 1.5..6 in 10 log-spaced scales, characterize on; one step is
bandpass -> scale-space maxima of all scales -> walk + characterize.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

LOG_PARAMS = dict(min_radius=1.5, max_radius=6.0, n_radii=10, log_spacing=True,
                  characterize=True)

WORKLOADS = {
    # name: (synth spec, frames, params)
    "log": ("log", 500, LOG_PARAMS),
    "full": ("full", 2000, dict(diameter=9, separation=7, minmass=400.0, characterize=True)),
    "small": ("small", 100, dict(diameter=9, separation=9, minmass=500.0, noise_size=1.0,
                                 smoothing_size=9, max_iterations=10, characterize=True)),
}


def _host_cpus():
    """(threads the CPU baseline uses, nproc, CPU model).  A one-GPU box
    allots 16 host cores to its GPU, whatever nproc shows, so the baseline
    runs one thread per allotted core."""
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(16, ncores)), os.cpu_count() or ncores, model


def _cpu_baseline(video_host, params, budget_s, gpu_rows):
    import numpy as np

    import oracle
    T, H, W = video_host.shape
    nthreads, nproc, model = _host_cpus()
    op = oracle.derive_trackpy(H, W, diameter=params["diameter"], separation=params["separation"],
                               minmass=params["minmass"])
    done = 0
    t0 = time.perf_counter()
    results = []
    while done < T and time.perf_counter() - t0 < budget_s:
        n = min(T - done, 4 * nthreads)
        rows, _ = oracle.detect_batch(video_host[done:done + n], op, characterize=True,
                                      nthreads=nthreads, cap_per_frame=8192)
        results.extend(rows)
        done += n
    dt = time.perf_counter() - t0
    # parity of the GPU rows on the sampled frames
    mismatched = 0
    fields = ["x", "y", "mass", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]
    for t in range(done):
        ref = np.sort(results[t], order="seed")
        got = gpu_rows[t]
        if len(ref) != len(got) or not np.array_equal(ref["seed"], got["seed"]) or not all(
                np.array_equal(ref[f].view(np.uint32), got[f].view(np.uint32)) for f in fields):
            mismatched += 1
    return {
        "value": round(done / dt, 3), "unit": "frames/s", "cores": nthreads, "kind": "port",
        "nproc": nproc, "cpu_model": model,
        "sample": f"first {done} frames of the workload, detect+characterize, oracle C code "
                  f"(-O2, one frame per thread) on {nthreads} host threads (the host cores a "
                  f"one-GPU box allots; nproc {nproc}), {dt:.1f} s",
    }, done, mismatched


def _fp64_error(video_host, params, gpu_rows, nframes=8):
    """Error of the f32 pipeline (the GPU rows, bit-identical to the oracle)
    against an FP64 recomputation: the bandpass in float64, then at the walk's
    final integer position (the 3 x 3 neighbour of the rounded centre whose
    float64 offset stays within the shift threshold) the float64 centroid and
    mass over the same disk.  Checker code of the CPU-baseline leg."""
    import numpy as np

    import oracle
    T, H, W = video_host.shape
    op = oracle.derive_trackpy(H, W, diameter=params["diameter"],
                               separation=params["separation"], minmass=params["minmass"])
    r = (params["diameter"] - 1) // 2
    ii, jj = np.mgrid[-r:r + 1, -r:r + 1]
    keep = (ii * ii + jj * jj) <= r * r
    di, dj = ii[keep], jj[keep]
    cerr, merr, unresolved = [], [], 0
    for t in range(min(nframes, T)):
        P64 = oracle.bandpass_f64(video_host[t], op)
        rows = gpu_rows[t]
        if len(rows) == 0:
            continue
        x = rows["x"].astype(np.float64)
        y = rows["y"].astype(np.float64)
        m32 = rows["mass"].astype(np.float64)
        best = np.full(len(rows), np.inf)
        bm = np.full(len(rows), np.nan)
        for du in (-1, 0, 1):
            for dv in (-1, 0, 1):
                pu = np.rint(x).astype(np.int64) + du
                pv = np.rint(y).astype(np.int64) + dv
                U = pu[:, None] + di[None]
                V = pv[:, None] + dj[None]
                ok = (U >= 0) & (U < H) & (V >= 0) & (V < W)
                d = np.where(ok, P64[np.clip(U, 0, H - 1), np.clip(V, 0, W - 1)], 0.0)
                m = d.sum(1)
                with np.errstate(invalid="ignore", divide="ignore"):
                    ox = (d * di[None]).sum(1) / m
                    oy = (d * dj[None]).sum(1) / m
                    e = np.maximum(np.abs(pu + ox - x), np.abs(pv + oy - y))
                    stay = (np.abs(ox) <= params.get("shift_threshold", 0.6)) & \
                        (np.abs(oy) <= params.get("shift_threshold", 0.6)) & (m > 0)
                better = stay & (e < best)
                best = np.where(better, e, best)
                bm = np.where(better, m, bm)
        found = np.isfinite(best)
        unresolved += int((~found).sum())
        cerr.append(best[found])
        merr.append(np.abs(m32[found] - bm[found]) / bm[found])
    c = np.concatenate(cerr) if cerr else np.zeros(0)
    m = np.concatenate(merr) if merr else np.zeros(0)
    return {"frames": min(nframes, T), "rows": int(len(c)), "unresolved_rows": unresolved,
            "centroid_max_px": float(c.max()) if len(c) else None,
            "centroid_p99_9_px": float(np.quantile(c, 0.999)) if len(c) else None,
            "mass_max_rel": float(m.max()) if len(m) else None,
            "mass_p99_9_rel": float(np.quantile(m, 0.999)) if len(m) else None,
            "reference": "float64 bandpass and centroid at the walk's final position "
                         "(the f32 rows are bit-identical to the oracle's)"}


def _e2e_oracle_check(video_host, params, data, names, search_range, memory, nframes=2):
    """The end-to-end rows of the first frames (detect, characterize,
    filter_close, doughnut) and their particle ids against the oracle's: a
    frame's ids depend only on it and the frames before it.  Checker code of
    the CPU-baseline leg."""
    import numpy as np

    import oracle
    from oracle import linker as OL
    T, H, W = video_host.shape
    d = params["diameter"]
    op = oracle.derive_trackpy(H, W, diameter=d, separation=params["separation"],
                               minmass=params["minmass"])
    ref = []
    for t in range(nframes):
        rows, _ = oracle.detect_frame(video_host[t], op, characterize=True)
        rows = np.sort(rows, order="seed")
        rows = rows[oracle.filter_close(rows, float(params["separation"]))]
        dn = oracle.doughnut(video_host[t], rows, d / 2.0, float(d), 0.0)
        ref.append(np.stack([np.full(len(rows), t, np.float32), rows["x"], rows["y"],
                             rows["mass"], rows["Rg"], rows["raw_mass"], rows["signal"],
                             rows["ecc"], dn[:, 0], dn[:, 1], dn[:, 2]], 1).astype(np.float32))
    ref = np.concatenate(ref, 0)
    got = data[data[:, list(names).index("frame")] < nframes]
    ids = got[:, list(names).index("particle")].astype(np.int64).tolist()
    same_rows = got.shape[0] == ref.shape[0] and np.array_equal(
        np.ascontiguousarray(got[:, :ref.shape[1]]).view(np.uint32), ref.view(np.uint32))
    same_ids = same_rows and ids == OL.link_frames(ref[:, :3], search_range, memory)
    return {"frames": nframes, "rows": int(ref.shape[0]), "rows_bit_identical": bool(same_rows),
            "particle_ids_identical": bool(same_ids)}


def _log_taps_per_pixel(params):
    """Sum over the scales of 2R+1, R = i32(4 sigma + 0.5), sigma = r / sqrt(2)
    (the radius scan of gpu_setup.rs:958-984 in f32)."""
    import numpy as np
    f = np.float32
    n = params["n_radii"]
    lo, hi = f(params["min_radius"]), f(params["max_radius"])
    if params["log_spacing"]:
        start, end = np.log10(lo), np.log10(hi)
        diff = f((end - start) / f(n - 1))
        radii = []
        for _ in range(n - 1):
            radii.append(f(10.0) ** start)
            start = f(start + diff)
        radii.append(f(10.0) ** start)
    else:
        diff = f((hi - lo) / f(n - 1))
        radii, start = [], lo
        for _ in range(n - 1):
            radii.append(start)
            start = f(start + diff)
        radii.append(start)
    return sum(2 * int(f(4.0) * f(r / np.sqrt(f(2.0))) + f(0.5)) + 1 for r in radii)


def _cpu_baseline_log(video_host, params, budget_s, gpu_rows):
    """The oracle's LoG frame (single-threaded C per frame) on host threads."""
    import threading

    import numpy as np

    import oracle
    T, H, W = video_host.shape
    nthreads, nproc, model = _host_cpus()
    op = oracle.derive_log(H, W, params["max_radius"])
    radii = oracle.log_radii(params["min_radius"], params["max_radius"], params["n_radii"],
                             params["log_spacing"])
    results = {}

    def work(t):
        # unpruned rows in seed order: the device path's output before the
        # host post-processing
        results[t], _ = oracle.log_detect_frame(video_host[t], op, radii, characterize=True,
                                                overlap_threshold=1.0)

    done = 0
    t0 = time.perf_counter()
    while done < T and time.perf_counter() - t0 < budget_s:
        n = min(T - done, nthreads)
        ths = [threading.Thread(target=work, args=(done + k,)) for k in range(n)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        done += n
    dt = time.perf_counter() - t0
    mismatched = 0
    fields = ["x", "y", "mass", "r", "max_intensity", "Rg", "raw_mass", "signal", "ecc", "count"]
    for t in range(done):
        ref, got = results[t], gpu_rows[t]
        if len(ref) != len(got) or not np.array_equal(ref["seed"], got["seed"]) or not all(
                np.array_equal(ref[f], got[f], equal_nan=True) for f in fields):
            mismatched += 1
    return {
        "value": round(done / dt, 3), "unit": "frames/s", "cores": nthreads, "kind": "port",
        "nproc": nproc, "cpu_model": model,
        "sample": f"first {done} frames of the workload, LoG detect+characterize (no pruning), "
                  f"oracle C code (-O2, one frame per thread) on {nthreads} host threads, "
                  f"{dt:.1f} s",
    }, done, mismatched


BATCH_SIZES = [(256, 256), (512, 384), (1024, 1024), (2048, 1024)]


def _row_checksums(torch, out, counts):
    """Per frame: sum over its rows of a hash of the row's 12 words (order
    independent), as int64."""
    n, cap, _ = out.shape
    mult = torch.tensor([0x9E3779B1 + 2 * k for k in range(12)], dtype=torch.int64,
                        device=out.device)
    res = []
    for a in range(0, n, 512):
        w = out[a:a + 512].to(torch.int64) & 0xFFFFFFFF
        h = ((w * mult) & 0xFFFFFFFFFFFF).sum(dim=2)                   # (chunk, cap)
        valid = torch.arange(cap, device=out.device)[None, :] < \
            counts[a:a + 512, None].to(torch.int64)
        res.append((h * valid).sum(dim=1))
    return torch.cat(res)


def run_batch_workload(args):
    import numpy as np
    import torch

    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    rng = np.random.default_rng(20261203 + 4000)
    nvid = 64
    size_idx = rng.integers(0, len(BATCH_SIZES), nvid)
    lengths = rng.integers(200, 1001, nvid)
    lengths = np.maximum(4, np.round(lengths * args.batch_scale).astype(int))
    params = dict(separation=7, minmass=400.0, characterize=True)
    p = gt.make_params(9, **params)
    print(f"[bench] rendering {nvid} videos, {int(lengths.sum())} frames", file=sys.stderr,
          flush=True)
    vids = []
    for v in range(nvid):
        H, W = BATCH_SIZES[size_idx[v]]
        sp = synth.spec("full", nrows=H, ncols=W, n_particles=H * W // 1000)
        vids.append(synth.video_torch(sp, int(lengths[v]), seed=20261203 + 4000 + v,
                                      device="cuda"))
        if v % 8 == 7:
            print(f"[bench] rendered {v + 1} videos", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    frames = [f for vid in vids for f in vid.unbind(0)]
    ids = [t for vid in vids for t in range(vid.shape[0])]
    npx = sum(f.numel() for f in frames)
    nf = len(frames)
    cap = 4096
    out_b = torch.empty((nf, cap, 12), dtype=torch.int32, device="cuda")
    cnt_b = torch.zeros((nf,), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    with gt.Batch(p, max_frame_pixels=2048 * 1024, max_launch_pixels=256 << 20,
                  max_slots=1024) as b:
        # autotune on the first 1 Gpx of frames (mixed sizes, several launches,
        # so that the lane count shows)
        acc, k = 0, 0
        while k < nf and acc < (1024 << 20):
            acc += frames[k].numel()
            k += 1
        print(f"[bench] autotuning on {k} frames", file=sys.stderr, flush=True)
        tuned_ms = b.tune(frames[:k], cap_per_frame=cap)
        sched = b.schedule()
        # the videos as 3-D tensors: their frames join the call in order
        for _ in range(max(1, args.warmup)):
            b.detect(vids, frame_ids=ids, out=out_b, counts=cnt_b, cap_per_frame=cap,
                     stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.detect(vids, frame_ids=ids, out=out_b, counts=cnt_b, cap_per_frame=cap,
                     stream=stream)
        torch.cuda.synchronize()
        dt_b = (time.perf_counter() - t0) / args.steps
        launches = b.schedule()["launches_last_call"]
        # the same launches without the graph, for comparison
        b.set_graph(False)
        b.detect(vids, frame_ids=ids, out=out_b, counts=cnt_b, cap_per_frame=cap, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.detect(vids, frame_ids=ids, out=out_b, counts=cnt_b, cap_per_frame=cap,
                     stream=stream)
        torch.cuda.synchronize()
        dt_direct = (time.perf_counter() - t0) / args.steps
    print(f"[bench] batched {nf / dt_b:.0f} frames/s ({launches} launches)", file=sys.stderr,
          flush=True)
    # separate one-size runs, one context per video
    out_s = torch.empty((nf, cap, 12), dtype=torch.int32, device="cuda")
    cnt_s = torch.zeros((nf,), dtype=torch.int32, device="cuda")
    ctxs = [gt.Context(p, vid.shape[1], vid.shape[2]) for vid in vids]
    offs = np.concatenate([[0], np.cumsum([vid.shape[0] for vid in vids])])

    def separate():
        for v, vid in enumerate(vids):
            ctxs[v].detect_device(vid, out=out_s[offs[v]:offs[v + 1]],
                                  counts=cnt_s[offs[v]:offs[v + 1]], cap_per_frame=cap,
                                  stream=stream)

    separate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        separate()
    torch.cuda.synchronize()
    dt_s = (time.perf_counter() - t0) / args.steps
    for c in ctxs:
        c.close()
    if int(cnt_b.max().item()) > cap:
        raise SystemExit("detection capacity exceeded")
    same_counts = bool(torch.equal(cnt_b, cnt_s))
    same_rows = bool(torch.equal(_row_checksums(torch, out_b, cnt_b),
                                 _row_checksums(torch, out_s, cnt_s)))
    fps = nf / dt_b
    achieved = 10.0 * npx / dt_b / 1e9
    res = {
        "metric": "frames/s (batched uneven videos, detect+characterize, u16)",
        "value": round(fps, 1), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt_b * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded, rendered on the GPU)",
        "config": {"workload": f"batch: {nvid} videos, sizes {BATCH_SIZES}, "
                               f"{nf} frames, {npx / 1e9:.2f} Gpx, 1 particle / 1000 px^2, "
                               "diameter=9, separation=7, minmass=400, characterize=True"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batched pipeline (10 B/px algorithmic)"},
        "pixels_per_s": round(npx / dt_b, 1),
        "separate_runs": {"frames_per_s": round(nf / dt_s, 1), "ms_per_pass": round(dt_s * 1e3, 3),
                          "speedup_of_batched": round(dt_s / dt_b, 3)},
        "schedule": sched, "tuned_sample_ms": round(tuned_ms, 3), "launches": launches,
        "direct_launches_frames_per_s": round(nf / dt_direct, 1),
        "parity": {"frames": nf, "counts_identical": same_counts,
                   "row_checksums_identical": same_rows},
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


# matrix-core multiply-adds of one forward pass of the centroid network:
# conv2 49 x 32 x 144, conv3 25 x 32 x 288 (conv1 and the head run on the VALU)
CNN_MFMA_FLOP = 2 * (49 * 32 * 144 + 25 * 32 * 288)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _match_rows(plain, learned, tol=1e-3):
    """Particles of `plain` (per-frame rows) found in `learned` (same seed)
    within tol px, and the row counts."""
    ok = n = extra = 0
    for a, b in zip(plain, learned):
        bs = {int(r["seed"]): r for r in b}
        for r in a:
            n += 1
            q = bs.pop(int(r["seed"]), None)
            if q is not None and abs(q["x"] - r["x"]) <= tol and abs(q["y"] - r["y"]) <= tol:
                ok += 1
        extra += len(bs)
    return ok, n, extra


def run_learned_workload(args):
    import numpy as np
    import torch

    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    s = synth.spec("full")
    H, W = s.nrows, s.ncols
    params = dict(separation=7, minmass=400.0, characterize=True)
    p = gt.make_params(9, **params)
    ctx = gt.Context(p, H, W)
    # training pairs from the same generator, seeds disjoint from the evaluation video
    want = args.learned_samples
    pchunks, tchunks, got, seed, nfr = [], [], 0, 20261203 + 9000, 0
    t0 = time.perf_counter()
    while got < want:
        v = synth.video_torch(s, 32, seed=seed, device="cuda")
        P, T, _ = ctx.training_samples(v)
        pchunks.append(P.clone())
        tchunks.append(T.clone())
        got += P.shape[0]
        nfr += 32
        seed += 1
        del v, P, T
    P = torch.cat(pchunks)[:want].contiguous()
    T = torch.cat(tchunks)[:want].contiguous()
    del pchunks, tchunks
    torch.cuda.synchronize()
    t_samples = time.perf_counter() - t0
    print(f"[bench] {P.shape[0]} training pairs from {nfr} frames in {t_samples:.1f} s "
          f"({float(T[:, 2].mean()):.3f} kept)", file=sys.stderr, flush=True)
    net = gt.CentroidNet(seed=7)
    t0 = time.perf_counter()
    losses = net.train(P, T, epochs=args.learned_epochs, batch=4096, lr=1e-3, seed=11)
    t_train = time.perf_counter() - t0
    steps_per_epoch = (P.shape[0] + 4095) // 4096
    print(f"[bench] trained {args.learned_epochs} epochs ({steps_per_epoch} steps each) in "
          f"{t_train:.1f} s, losses {losses}", file=sys.stderr, flush=True)

    # the network's forward kernel alone over stored patches (MFMA roofline)
    nfw = min(P.shape[0], 1 << 20)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    net.forward(P[:nfw])
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        net.forward(P[:nfw])
    e1.record()
    torch.cuda.synchronize()
    fw_s = e0.elapsed_time(e1) * 1e-3 / 5
    fw_tflops = nfw * CNN_MFMA_FLOP / fw_s / 1e12

    T_eval = args.frames or 2000
    frames = synth.video_torch(s, T_eval, seed=20261203 + 2000, device="cuda")
    cap = 8192
    out_p = torch.empty((T_eval, cap, 12), dtype=torch.int32, device="cuda")
    cnt_p = torch.zeros((T_eval,), dtype=torch.int32, device="cuda")
    out_l = torch.empty_like(out_p)
    cnt_l = torch.zeros_like(cnt_p)
    stream = torch.cuda.current_stream().cuda_stream

    def timed(out, cnt):
        for _ in range(args.warmup):
            ctx.detect_device(frames, out=out, counts=cnt, cap_per_frame=cap, stream=stream)
        torch.cuda.synchronize()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.detect_device(frames, out=out, counts=cnt, cap_per_frame=cap, stream=stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = ctx.stage_times()
        ctx.set_timing(False)
        return T_eval * args.steps / dt, st

    fps_plain, st_plain = timed(out_p, cnt_p)
    nchk = min(T_eval, 256)
    rows_p = gt.detections_to_numpy(out_p[:nchk], cnt_p[:nchk])
    scopes = {}
    for scope in ("all", "moving"):      # "moving" last: its rows stay in out_l
        ctx.set_learned(net, scope=scope)
        ctx.learned_stats(reset=True)
        fps_l, st_l = timed(out_l, cnt_l)
        acc, cands = ctx.learned_stats(reset=True)
        ctx.set_learned(None)
        rows_l = gt.detections_to_numpy(out_l[:nchk], cnt_l[:nchk])
        ok, n, extra = _match_rows(rows_p, rows_l)
        scopes[scope] = {
            "frames_per_s": round(fps_l, 1), "speedup_vs_plain": round(fps_l / fps_plain, 4),
            "stage_ms_per_step": [round(x, 3) for x in st_l[0]],
            "proposals": {"accepted": acc, "listed": cands,
                          "accepted_fraction": round(acc / max(cands, 1), 4),
                          "listed_per_frame": round(cands / (T_eval * (args.steps + args.warmup)), 1)},
            "accuracy": {"frames": nchk, "particles": n, "within_1e-3_px": ok,
                         "fraction": round(ok / max(n, 1), 6), "extra_rows": extra}}
        print(f"[bench] learned ({scope}) {fps_l:.0f} frames/s vs plain {fps_plain:.0f}; "
              f"stages {st_l[0]}; {ok}/{n} within 1e-3 px", file=sys.stderr, flush=True)
    fps_learned = scopes["moving"]["frames_per_s"]
    cpu = None
    if not args.no_cpu:
        host = frames[:min(T_eval, 2048)].cpu().numpy().view(np.uint16)
        rows = gt.detections_to_numpy(out_p[:host.shape[0]], cnt_p[:host.shape[0]])
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu, _, _ = _cpu_baseline(host, dict(diameter=9, **params), args.cpu_seconds, rows)
    ctx.close()
    res = {
        "metric": "frames/s (learned centroid proposal for moving walks, detect+characterize, "
                  "1024^2 u16)",
        "value": round(fps_learned, 1), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(T_eval / fps_learned * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded, rendered on the GPU)",
        "config": {"workload": f"learned: {T_eval} frames {H}x{W} u16, {s.n_particles} "
                               "particles/frame, diameter=9, separation=7, minmass=400, "
                               "characterize=True; network trained on "
                               f"{P.shape[0]} pairs from {nfr} frames of disjoint seeds"},
        "plain_frames_per_s": round(fps_plain, 1),
        "speedup_vs_plain": round(fps_learned / fps_plain, 4),
        "plain_stage_ms_per_step": [round(x, 3) for x in st_plain[0]],
        "scopes": scopes,
        "accuracy": scopes["moving"]["accuracy"],
        "training": {"pairs": int(P.shape[0]), "frames": nfr, "epochs": args.learned_epochs,
                     "batch": 4096, "lr": 1e-3, "steps": steps_per_epoch * args.learned_epochs,
                     "seconds": round(t_train, 2), "sample_seconds": round(t_samples, 2),
                     "loss_per_epoch": [round(float(x), 5) for x in losses]},
        "roofline": {"bound": "mfma", "achieved": round(fw_tflops, 2),
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(fw_tflops / MFMA_BF16_PEAK_TFLOPS, 5), "traffic": None,
                     "kernel": "k_cnn_forward", "patches": nfw,
                     "alg_flop_per_patch": CNN_MFMA_FLOP,
                     "launch_ms": round(fw_s * 1e3, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    net.close()


def _copy_peak(torch):
    """Device copy bandwidth (GB/s, read + write bytes) over 1 GiB: the
    faster of torch's copy and elementwise-multiply kernels, best of three
    each, timed with events on the current stream."""
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    best = float("inf")
    for op in (lambda: dst.copy_(src), lambda: torch.mul(src, 2.0, out=dst)):
        op()
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            op()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
    del src, dst
    return 2.0 * (1 << 30) / best / 1e9


def _learned_summary(frames, params, nframes, steps):
    """The learned centroid proposal on the main workload's frames: a
    network trained on the GPU (2M pairs from frames of the same generator
    with disjoint seeds, Adam lr 1e-3, batch 4096, 4 epochs), then plain and
    learned detection of the same frames (speed-up, accuracy against the
    plain walk) and the network kernel's matrix-core roofline."""
    import torch

    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    s = synth.spec("full")
    H, W = frames.shape[1], frames.shape[2]
    p = gt.make_params(params["diameter"], **{k: v for k, v in params.items() if k != "diameter"})
    ctx = gt.Context(p, H, W)
    want, got, seed, nfr = 2_000_000, 0, 20261203 + 9000, 0
    pch, tch = [], []
    t0 = time.perf_counter()
    while got < want:
        v = synth.video_torch(s, 32, seed=seed, device="cuda")
        P, T, _ = ctx.training_samples(v)
        pch.append(P.clone())
        tch.append(T.clone())
        got += P.shape[0]
        nfr += 32
        seed += 1
        del v, P, T
    P = torch.cat(pch)[:want].contiguous()
    T = torch.cat(tch)[:want].contiguous()
    del pch, tch
    net = gt.CentroidNet(seed=7)
    losses = net.train(P, T, epochs=4, batch=4096, lr=1e-3, seed=11)
    t_train = time.perf_counter() - t0
    nfw = min(P.shape[0], 1 << 20)
    net.forward(P[:nfw])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        net.forward(P[:nfw])
    e1.record()
    torch.cuda.synchronize()
    fw_s = e0.elapsed_time(e1) * 1e-3 / 3
    del P, T
    fr = frames[:nframes]
    cap = 8192
    out_p = torch.empty((nframes, cap, 12), dtype=torch.int32, device="cuda")
    cnt_p = torch.zeros((nframes,), dtype=torch.int32, device="cuda")
    out_l, cnt_l = torch.empty_like(out_p), torch.zeros_like(cnt_p)
    stream = torch.cuda.current_stream().cuda_stream

    def timed(out, cnt):
        ctx.detect_device(fr, out=out, counts=cnt, cap_per_frame=cap, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.detect_device(fr, out=out, counts=cnt, cap_per_frame=cap, stream=stream)
        torch.cuda.synchronize()
        return nframes * steps / (time.perf_counter() - t0)

    fps_plain = timed(out_p, cnt_p)
    ctx.set_learned(net, scope="moving")
    fps_l = timed(out_l, cnt_l)
    acc, cands = ctx.learned_stats(reset=True)
    ctx.set_learned(None)
    nchk = min(nframes, 256)
    ok, n, extra = _match_rows(gt.detections_to_numpy(out_p[:nchk], cnt_p[:nchk]),
                               gt.detections_to_numpy(out_l[:nchk], cnt_l[:nchk]))
    ctx.close()
    net.close()
    tflops = nfw * CNN_MFMA_FLOP / fw_s / 1e12
    return {"frames": nframes, "scope": "moving", "plain_frames_per_s": round(fps_plain, 1),
            "learned_frames_per_s": round(fps_l, 1),
            "speedup_vs_plain": round(fps_l / fps_plain, 4),
            "accuracy": {"frames": nchk, "particles": n, "within_1e-3_px_of_full_walk": ok,
                         "fraction": round(ok / max(n, 1), 6), "extra_rows": extra},
            "proposals_accepted_fraction": round(acc / max(cands, 1), 4),
            "training": {"pairs": want, "frames": nfr, "epochs": 4, "batch": 4096, "lr": 1e-3,
                         "seconds_with_sampling": round(t_train, 2),
                         "final_loss": round(float(losses[-1]), 5)},
            "roofline": {"bound": "mfma", "kernel": "k_cnn_forward",
                         "achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5)}}


def _log_summary(nframes, steps):
    """The LoG workload at its shape (2048 x 2048 f32 frames of its
    generator, radii 1.5..6 in 10 log-spaced scales, characterize on), with
    threshold_rel off and at 0.05: frames/s and the scale-space kernel's FP32
    vector roofline."""
    import torch

    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    s = synth.spec("log")
    H, W = s.nrows, s.ncols
    frames = synth.video_torch(s, nframes, seed=20261203 + 3000, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    res = {"frames": nframes, "shape": [H, W], "dtype": "f32"}
    taps = _log_taps_per_pixel(LOG_PARAMS)
    for rel in (0.0, 0.05):
        kw = {k: v for k, v in LOG_PARAMS.items() if k not in ("min_radius", "max_radius")}
        p = gt.make_log_params(LOG_PARAMS["min_radius"], LOG_PARAMS["max_radius"],
                               threshold_rel=rel, **kw)
        cap = (H * W) // 8 if rel == 0.0 else 16384
        with gt.Context(p, H, W) as ctx:
            out = torch.empty((nframes, cap, 12), dtype=torch.int32, device="cuda")
            cnt = torch.zeros((nframes,), dtype=torch.int32, device="cuda")
            ctx.detect_device(frames, out=out, counts=cnt, cap_per_frame=cap, stream=stream)
            torch.cuda.synchronize()
            ctx.set_timing(True)
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.detect_device(frames, out=out, counts=cnt, cap_per_frame=cap, stream=stream)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            st, nl, _ = ctx.stage_times()
            ctx.set_timing(False)
        fps = nframes * steps / dt
        flop = 8.0 * taps * H * W * nframes
        tk = st[1] * 1e-3 if st else float("nan")
        res[f"threshold_rel_{rel}"] = {
            "frames_per_s": round(fps, 1),
            "mean_rows_per_frame": round(float(cnt.float().mean().item()), 1),
            "stage_ms_per_step": [round(x, 3) for x in st],
            "pipeline_frac": round((4 + 8 + 8 * LOG_PARAMS["n_radii"]) * H * W * fps
                                   / (HBM_PEAK_GBS * 1e9), 4),
            "log_tile_valu": {"achieved_tflops": round(flop / tk / 1e12, 2),
                              "peak_tflops": 78.6, "frac": round(flop / tk / 1e12 / 78.6, 4)}}
        del out, cnt
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="full",
                    choices=sorted(WORKLOADS) + ["batch", "learned"])
    ap.add_argument("--frames", type=int, default=0, help="override the frame count")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="batch workload: scale of the video lengths")
    ap.add_argument("--e2e-frames", type=int, default=1024,
                    help="frames of the end-to-end run (host frames in, DataFrame columns out, "
                         "with linking; plain workloads only; 0 = skip)")
    ap.add_argument("--learned-samples", type=int, default=2_000_000,
                    help="learned workload: training pairs")
    ap.add_argument("--learned-epochs", type=int, default=4,
                    help="learned workload: training epochs")
    ap.add_argument("--no-lanes", action="store_true",
                    help="skip timing the steps with the chunks alternating between two "
                         "streams, and with each chunk's tail on a second stream (done after "
                         "the timed region by default; reported as two_lanes and overlap)")
    ap.add_argument("--no-learned", action="store_true",
                    help="full workload: skip the learned-proposal summary")
    ap.add_argument("--no-log", action="store_true",
                    help="full workload: skip the LoG-workload summary")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device-to-device copy (a known-size kernel that "
                         "calibrates rocprof's FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus above 1 is not supported (frames and videos "
              "shard across GPUs with no exchange, so N GPUs run N independent pipelines)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        run_batch_workload(args)
        return
    if args.workload == "learned":
        run_learned_workload(args)
        return

    import numpy as np
    import torch

    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth

    spec_name, T, params = WORKLOADS[args.workload]
    if args.frames:
        T = args.frames
    s = synth.spec(spec_name)
    H, W = s.nrows, s.ncols
    if args.copy_probe:
        # filled now, read after the frames are rendered: by then it is out of
        # every cache, so the probe's reads come from HBM
        probe_src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")    # 1 GiB
    print(f"[bench] rendering {T} frames {H}x{W} ({spec_name}) on the GPU", file=sys.stderr,
          flush=True)
    frames = synth.video_torch(s, T, seed=20261203 + 2000, device="cuda")
    torch.cuda.synchronize()

    if args.copy_probe:
        dst = torch.empty_like(probe_src)
        torch.mul(probe_src, 2.0, out=dst)   # one elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del probe_src, dst

    is_log = args.workload == "log"
    if is_log:
        p = gt.make_log_params(params["min_radius"], params["max_radius"],
                               **{k: v for k, v in params.items()
                                  if k not in ("min_radius", "max_radius")})
    else:
        p = gt.make_params(params["diameter"],
                           **{k: v for k, v in params.items() if k != "diameter"})
    ctx = gt.Context(p, H, W)
    # detections per frame: ~1000 on the plain workloads; LoG keeps every
    # scale-space maximum of a noisy frame (no threshold), ~7% of the pixels
    cap = (H * W) // 8 if is_log else 8192
    out = torch.empty((T, cap, 12), dtype=torch.int32, device="cuda")
    counts = torch.zeros((T,), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        ctx.detect_device(frames, out=out, counts=counts, cap_per_frame=cap, stream=stream)

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stage_ms, nlaunch, chunk = ctx.stage_times()   # device time of the last step
    ctx.set_timing(False)
    fps = T * args.steps / dt
    print(f"[bench] {fps:.0f} frames/s, stages (last step, ms) {stage_ms}", file=sys.stderr,
          flush=True)
    if int(counts.max().item()) > cap:
        raise SystemExit(f"detection capacity {cap} exceeded: {int(counts.max().item())}")

    # roofline of the dominant kernel; algorithmic bytes per pixel:
    # plain: bandpass 2 (u16 in) + 4 (f32 out); detect (local max + walk +
    # characterize) 4 (f32 in); finish (walks handed over + counts) ~0.
    # LoG: bandpass 4 + 4; log_tile 4 (raw f32 read once; the scale planes
    # stay in LDS); walk ~0.
    if is_log:
        names = ["bandpass", "log_tile", "walk"]
        alg_bpp = {"bandpass": 8.0, "log_tile": 4.0, "walk": 0.0}
        pipe_bpp = 4 + 8 + 8 * params["n_radii"]
    else:
        names = ["bandpass", "detect", "finish"]
        alg_bpp = {"bandpass": 6.0, "detect": 4.0, "finish": 0.0}
        pipe_bpp = 10
    k = int(np.argmax(stage_ms)) if stage_ms else 0
    kname = names[k]
    # per launch: the step's T frames over nlaunch launches (the last one
    # may hold fewer than `chunk`), so bytes and time are both averages
    fpl = T / nlaunch if stage_ms else float(chunk)
    launch_ms = stage_ms[k] / nlaunch if stage_ms else float("nan")
    bytes_launch = alg_bpp[kname] * H * W * fpl
    achieved = bytes_launch / (launch_ms * 1e-3) / 1e9 if stage_ms and alg_bpp[kname] else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{kname}_{args.workload}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        traffic = round(tj["bytes_per_frame"] * fpl)
    copy_gbs = _copy_peak(torch)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname, "launch_ms": round(launch_ms, 4),
            "frames_per_launch": round(fpl, 2), "launches_per_step": nlaunch,
            "alg_bytes_per_frame": alg_bpp[kname] * H * W,
            "stage_ms_per_step": [round(x, 3) for x in stage_ms],
            "copy_peak_gbs": round(copy_gbs, 1),
            "frac_of_copy_peak": round(achieved / copy_gbs, 4),
            "pipeline_alg_bytes_per_frame": pipe_bpp * H * W,
            "pipeline_frac": round(pipe_bpp * H * W * fps / (HBM_PEAK_GBS * 1e9), 4),
            "pipeline_frac_of_copy_peak": round(pipe_bpp * H * W * fps / (copy_gbs * 1e9), 4)}
    # the same kernel's mean dispatch time in the committed rocprof summary of
    # this command (profiles/), for comparison with the event-timed figure
    kf = os.path.join(ROOT, "profiles", f"kernel_times_{args.workload}.json")
    if os.path.exists(kf) and stage_ms:
        with open(kf) as f:
            kj = json.load(f)
        if kname in kj.get("kernels", {}):
            kk = kj["kernels"][kname]
            r_ach = alg_bpp[kname] * H * W * kk["frames_per_dispatch"] / (kk["avg_us"] * 1e-6) / 1e9
            roof["rocprof"] = {"kernel": kk["name"], "avg_us": kk["avg_us"],
                               "frames_per_dispatch": kk["frames_per_dispatch"],
                               "achieved": round(r_ach, 1),
                               "frac": round(r_ach / HBM_PEAK_GBS, 4),
                               "source": os.path.relpath(kf, ROOT)}
    if is_log and stage_ms:
        # the scale-space kernel is FP32-FMA bound: 4 separable passes of
        # 2R+1 taps per pixel per scale (8 (2R+1) flop), against the vector
        # FP32 peak 256 CU x 64 lanes x 2 flop x 2.4 GHz = 78.6 TFLOP/s
        taps = _log_taps_per_pixel(params)
        flop_frame = 8.0 * taps * H * W
        tk = stage_ms[1] / nlaunch * 1e-3
        roof["valu"] = {"kernel": "log_tile", "alg_flop_per_frame": flop_frame,
                        "achieved_tflops": round(flop_frame * chunk / tk / 1e12, 2),
                        "peak_tflops": 78.6,
                        "frac": round(flop_frame * chunk / tk / 1e12 / 78.6, 4)}

    lanes2 = overlap = None
    if not is_log and not args.no_lanes and args.workload in ("full", "small"):
        # the same steps with the chunks alternating between two streams
        # (outside the timed region above; kernel spans then overlap, so the
        # roofline is taken from the one-lane run)
        ref_counts = counts.clone()
        ref_sums = _row_checksums(torch, out, counts)
        def timed_mode():
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            same = bool(torch.equal(ref_counts, counts)) and \
                bool(torch.equal(ref_sums, _row_checksums(torch, out, counts)))
            return {"value": round(T * args.steps / dt, 1), "unit": "frames/s",
                    "identical_to_one_lane": same}

        ctx.set_lanes(2)
        lanes2 = timed_mode()
        ctx.set_lanes(1)
        ctx.set_overlap(True)
        overlap = timed_mode()
        ctx.set_overlap(False)
        print(f"[bench] two lanes {lanes2}, overlap {overlap}", file=sys.stderr, flush=True)

    e2e = None
    if not is_log and args.e2e_frames > 0:
        # whole-video path: host frames staged through pinned memory, detection
        # on the GPU, read-back, and filter_close + doughnut + linking
        # (search_range 5, memory 2) on the host thread, overlapped with the
        # next chunk's detection
        n = min(T, args.e2e_frames)
        host = frames[:n].cpu().numpy().view(np.uint16)
        kw = {k: v for k, v in params.items() if k != "diameter"}
        p2 = gt.make_params(params["diameter"], search_range=5.0, memory=2, **kw)
        with gt.Context(p2, H, W) as ctx2:
            ctx2.execute(host[:min(n, 8)])           # warm-up (allocations, first launch)
            t0 = time.perf_counter()
            data, names, _ = ctx2.execute(host)
            dt2 = time.perf_counter() - t0
        e2e = {"frames": n, "value": round(n / dt2, 1), "unit": "frames/s",
               "rows": int(data.shape[0]), "columns": list(names),
               "includes": "host frames -> pinned staging -> H2D -> detection -> D2H -> "
                           "filter_close, doughnut, linking (search_range 5, memory 2)"}
        e2e_sample = (host[:2].copy(), data, list(names))
        print(f"[bench] end to end {e2e['value']} frames/s", file=sys.stderr, flush=True)
        # the first frames (at most 256) from a multi-page u16 TIFF, read into
        # the pinned staging by the library's reader threads (the file is in
        # the page cache after it is written); every column but duration,
        # which depends on later frames, equals the array run's
        import tempfile
        nt = min(n, 256)
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "bench.tif")
            gt.to_tiff(path, host[:nt].astype(np.float32))
            with gt.VideoFile(path) as vf, gt.Context(p2, H, W) as ctx3:
                ctx3.execute(vf, keys=range(min(nt, 8)))
                t0 = time.perf_counter()
                data3, _, _ = ctx3.execute(vf)
                dt3 = time.perf_counter() - t0
        head = data[data[:, 0] < nt]
        same = data3.shape == head.shape and np.array_equal(
            np.ascontiguousarray(data3[:, :-1]).view(np.uint32),
            np.ascontiguousarray(head[:, :-1]).view(np.uint32))
        e2e["from_tiff"] = {"frames": nt, "value": round(nt / dt3, 1), "unit": "frames/s",
                            "identical_to_array_run": bool(same)}
        print(f"[bench] end to end from a TIFF file {e2e['from_tiff']['value']} frames/s "
              f"(identical: {same})", file=sys.stderr, flush=True)

    learned = log_sum = None
    if args.workload == "full" and not args.no_learned:
        print("[bench] learned-proposal summary", file=sys.stderr, flush=True)
        learned = _learned_summary(frames, params, min(T, 512), 3)
        print(f"[bench] learned {learned['speedup_vs_plain']}x plain, "
              f"{learned['accuracy']['fraction']} within 1e-3 px", file=sys.stderr, flush=True)
    if args.workload == "full" and not args.no_log:
        print("[bench] LoG-workload summary", file=sys.stderr, flush=True)
        log_sum = _log_summary(64, 3)
        print(f"[bench] LoG {log_sum}", file=sys.stderr, flush=True)

    cpu = None
    parity = None
    fp64 = None
    if not args.no_cpu:
        if is_log:
            rows = gt.detections_to_numpy(out[:64], counts[:64])
            host = frames[:64].cpu().numpy()
            print("[bench] CPU baseline", file=sys.stderr, flush=True)
            cpu, nchk, bad = _cpu_baseline_log(host, params, args.cpu_seconds, rows)
        else:
            # enough frames for the time budget; the parity check covers all
            # the frames the CPU finishes
            nh = min(T, 2048)
            rows = gt.detections_to_numpy(out[:nh], counts[:nh])
            host = frames[:nh].cpu().numpy().view(np.uint16)
            print("[bench] CPU baseline", file=sys.stderr, flush=True)
            cpu, nchk, bad = _cpu_baseline(host, params, args.cpu_seconds, rows)
            fp64 = _fp64_error(host, params, rows)
            if e2e is not None:
                e2e["oracle_check"] = _e2e_oracle_check(e2e_sample[0], params, e2e_sample[1],
                                                        e2e_sample[2], 5.0, 2)
                print(f"[bench] end to end against the oracle {e2e['oracle_check']}",
                      file=sys.stderr, flush=True)
        parity = {"frames_checked": nchk, "frames_bit_identical": nchk - bad}

    res = {
        "metric": "frames/s (detect+characterize, 1024^2 u16)" if args.workload == "full"
        else (f"frames/s (LoG detect+characterize, {H}x{W} f32)" if is_log
              else f"frames/s (detect+characterize, {H}x{W} u16)"),
        "value": round(fps, 1), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded, rendered on the GPU)",
        "config": {"workload": f"{args.workload}: {T} frames {H}x{W} "
                               f"{'f32' if is_log else 'u16'}, "
                               f"{s.n_particles} particles/frame, " +
                               ", ".join(f"{k}={v}" for k, v in params.items())},
        "roofline": roof, "cpu_baseline": cpu, "parity": parity, "end_to_end": e2e,
        "two_lanes": lanes2, "overlap": overlap,
        "mean_detections_per_frame": round(float(counts.float().mean().item()), 1),
        "fp64_error": fp64, "learned": learned, "log": log_sum,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
