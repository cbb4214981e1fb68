"""Learned centroid proposal on the GPU.

* The forward kernel, gradients and the Adam update are checked against the
  CPU oracle (oracle/cnn.py, pinned by torch autograd, finite differences
  and torch.optim.Adam).  Tolerances are those of bf16 activations with
  fp32 accumulation in a different order; Adam is bit-exact given the same
  gradient.
* Training samples (candidates, patches, walked targets and labels) are
  checked bit for bit against the oracle's bandpass, local max and walk.
* A trained network is checked to beat the method without it (the
  candidate pixel itself as the proposal).
* The learned detection path is checked against the plain walk: bit-identical
  when every proposal falls back, and equal centroids (1e-3 px) on at least
  99.9% of particles when proposals are used.
"""
import numpy as np
import pytest

from oracle import cnn as OC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev(gpu_ok):
    import torch
    return torch


def _params(seed):
    p = OC.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for k in ("b1", "b2", "b3", "bfc"):
        o, s = OC.OFF[k]
        p[o:o + int(np.prod(s))] = rng.uniform(-0.1, 0.1, int(np.prod(s)))
    return p


def _patch_tensor(torch, x):
    """(n, 11, 11) bf16 values -> device bfloat16 (n, 128)."""
    n = x.shape[0]
    buf = np.zeros((n, 128), np.float32)
    buf[:, :121] = x.reshape(n, 121)
    return torch.from_numpy(buf).to(torch.bfloat16).cuda()


def _samples(seed, n):
    rng = np.random.default_rng(seed)
    # Gaussian blobs off-centre plus noise: the kind of patch the detector sees
    yy, xx = np.mgrid[-5:6, -5:6]
    dy, dx = rng.uniform(-1.5, 1.5, (2, n))
    s = rng.uniform(1.2, 2.0, n)
    raw = np.exp(-((yy[None] - dy[:, None, None]) ** 2 + (xx[None] - dx[:, None, None]) ** 2)
                 / (2 * s[:, None, None] ** 2)) * rng.uniform(50, 500, n)[:, None, None]
    raw = raw + rng.normal(0, 5, raw.shape)
    x = OC.normalize_patches(raw.reshape(n, 121).astype(np.float32))
    t = np.stack([dy, dx, (rng.random(n) < 0.7).astype(np.float64), np.zeros(n)], 1)
    return x, t.astype(np.float32)


def test_forward_matches_oracle(torch_dev):
    import gpu_tracking_amd as gt
    p = _params(11)
    x, _ = _samples(11, 301)
    with gt.CentroidNet(params=p) as net:
        got = net.forward(_patch_tensor(torch_dev, x)).cpu().numpy()
    ref = OC.forward(p, x, rounding=True)
    assert np.all(got[:, 3] == 0)
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got[:, :3] - ref) / scale
    # bf16 activations: a rounding flip moves an output by ~2^-8 of its scale
    assert err.max() < 2e-2, err.max()
    assert np.median(err) < 2e-3


def test_gradient_matches_oracle(torch_dev):
    import gpu_tracking_amd as gt
    p = _params(12)
    x, t = _samples(12, 37)          # not a multiple of the four samples a block holds
    with gt.CentroidNet(params=p) as net:
        g, L = net.gradient(_patch_tensor(torch_dev, x), torch_dev.from_numpy(t).cuda())
    ref, Lref = OC.gradients(p, x, t, rounding=True)
    assert abs(L - Lref) <= 1e-3 * abs(Lref)
    for k, (o, s) in OC.OFF.items():
        n = int(np.prod(s))
        a, b = g[o:o + n].astype(np.float64), ref[o:o + n]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 3e-2, (k, rel)


def test_gradient_of_a_permuted_subset(torch_dev):
    import gpu_tracking_amd as gt
    p = _params(13)
    x, t = _samples(13, 90)
    idx = np.random.default_rng(13).permutation(90)[:41]
    with gt.CentroidNet(params=p) as net:
        g, L = net.gradient(_patch_tensor(torch_dev, x), torch_dev.from_numpy(t).cuda(),
                            idx=torch_dev.from_numpy(idx.astype(np.int32)))
    ref, Lref = OC.gradients(p, x[idx], t[idx], rounding=True)
    assert abs(L - Lref) <= 1e-3 * abs(Lref)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 3e-2


def test_adam_step_matches_oracle(torch_dev):
    import gpu_tracking_amd as gt
    p = _params(14)
    x, t = _samples(14, 64)
    P, T = _patch_tensor(torch_dev, x), torch_dev.from_numpy(t).cuda()
    with gt.CentroidNet(params=p) as net:
        g, _ = net.gradient(P, T)
        assert net.adam_steps == 0
        net.step(P, T, lr=1e-3)
        assert net.adam_steps == 1
        got = net.params
    want, _, _ = OC.adam_step(p, g, np.zeros_like(p), np.zeros_like(p), 1, lr=1e-3)
    np.testing.assert_array_equal(got, want)


def _frames(n, seed, H=256, W=256, density=600):
    from gpu_tracking_amd import synth
    s = synth.spec("full", nrows=H, ncols=W, n_particles=max(8, H * W // density))
    return synth.video_numpy(s, n, seed=seed)


PRM = dict(separation=7, minmass=400.0)


def test_training_samples_match_oracle(torch_dev):
    import oracle
    import gpu_tracking_amd as gt
    v = _frames(2, seed=21)
    op = oracle.derive_trackpy(v.shape[1], v.shape[2], diameter=9, separation=7, minmass=400.0)
    with gt.Context(gt.make_params(9, **PRM), v.shape[1], v.shape[2]) as ctx:
        for f in range(v.shape[0]):
            frame = torch_dev.from_numpy(v[f:f + 1].view(np.int16)).cuda()
            patches, targets, total = ctx.training_samples(frame)
            torch_dev.cuda.synchronize()
            tg = targets.cpu().numpy()
            assert len(tg) == total > 0
            seeds = tg[:, 3].view(np.uint32)
            proc = oracle.bandpass(v[f], op)
            cand = oracle.local_max(proc, op)
            assert np.array_equal(np.sort(seeds), np.sort(cand))
            order = np.argsort(seeds)
            seeds, tg = seeds[order], tg[order]
            pb = patches.view(torch_dev.int16).cpu().numpy()[order].view(np.uint16)
            cu, cv = seeds // v.shape[2], seeds % v.shape[2]
            want = OC.normalize_patches(OC.gather_patches(proc, np.stack([cu, cv], 1)))
            want_bits = (want.reshape(-1, 121).view(np.uint32) >> 16).astype(np.uint16)
            assert np.array_equal(pb[:, :121], want_bits)
            assert np.all(pb[:, 121:] == 0)
            # every candidate's walk end (minmass 0 keeps all); a walk
            # converged within 10 steps iff 11 steps end at the same place
            op0 = oracle.derive_trackpy(v.shape[1], v.shape[2], diameter=9, separation=7)
            op1 = oracle.derive_trackpy(v.shape[1], v.shape[2], diameter=9, separation=7,
                                        max_iterations=11)
            end10 = {int(r["seed"]): (r["x"], r["y"]) for r in oracle.walk(proc, op0, cand)}
            end11 = {int(r["seed"]): (r["x"], r["y"]) for r in oracle.walk(proc, op1, cand)}
            assert len(end10) == len(cand)
            for s, t, u, w in zip(seeds, tg, cu, cv):
                e10, e11 = end10[int(s)], end11[int(s)]
                conv = np.isfinite(e10[0]) and np.isfinite(e10[1]) and \
                    e10[0] == e11[0] and e10[1] == e11[1]
                assert t[2] == (1.0 if conv else 0.0)
                if conv:
                    assert np.float32(u) + t[0] == e10[0] and np.float32(w) + t[1] == e10[1]
                else:
                    assert t[0] == 0.0 and t[1] == 0.0


@pytest.fixture(scope="module")
def trained(torch_dev):
    """A network trained on samples from frames with seeds 100..; returns
    (net, per-epoch losses)."""
    import gpu_tracking_amd as gt
    v = _frames(24, seed=100, H=384, W=384)
    with gt.Context(gt.make_params(9, **PRM), v.shape[1], v.shape[2]) as ctx:
        P, T, _ = ctx.training_samples(torch_dev.from_numpy(v.view(np.int16)).cuda())
    net = gt.CentroidNet(seed=3)
    losses = net.train(P, T, epochs=12, batch=512, lr=2e-3, seed=5)
    yield net, losses
    net.close()


def test_training_beats_the_method_without_it(torch_dev, trained):
    import gpu_tracking_amd as gt
    net, losses = trained
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0], losses
    # held-out frames (other seed)
    v = _frames(6, seed=500, H=384, W=384)
    with gt.Context(gt.make_params(9, **PRM), v.shape[1], v.shape[2]) as ctx:
        P, T, _ = ctx.training_samples(torch_dev.from_numpy(v.view(np.int16)).cuda())
    o = net.forward(P).cpu().numpy()
    t = T.cpu().numpy()
    pos = t[:, 2] == 1.0
    err_net = np.hypot(o[pos, 0] - t[pos, 0], o[pos, 1] - t[pos, 1])
    err_none = np.hypot(t[pos, 0], t[pos, 1])        # proposal = the candidate pixel
    print(f"\nmean end-position error: network {err_net.mean():.4f} px, "
          f"candidate pixel {err_none.mean():.4f} px; "
          f"confidence accuracy {np.mean((o[:, 2] >= 0) == pos):.4f}")
    assert err_net.mean() < 0.5 * err_none.mean()
    assert np.mean((o[:, 2] >= 0) == pos) > 0.9


def _rows(out, counts):
    import gpu_tracking_amd as gt
    return gt.detections_to_numpy(out, counts)


@pytest.mark.parametrize("scope", ["moving", "all"])
def test_learned_path_without_proposals_is_the_plain_path(torch_dev, trained, scope):
    import gpu_tracking_amd as gt
    net, _ = trained
    v = _frames(4, seed=700, H=384, W=384)
    frames = torch_dev.from_numpy(v.view(np.int16)).cuda()
    p = gt.make_params(9, characterize=True, **PRM)
    with gt.Context(p, v.shape[1], v.shape[2]) as ctx:
        plain = _rows(*ctx.detect_device(frames))
        ctx.set_learned(net, conf_logit=float("inf"), scope=scope)
        fallback = _rows(*ctx.detect_device(frames))
        acc, cands = ctx.learned_stats()
    assert acc == 0 and cands > 0
    for a, b in zip(plain, fallback):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("scope", ["moving", "all"])
def test_learned_path_matches_the_full_walk(torch_dev, trained, scope):
    import gpu_tracking_amd as gt
    net, _ = trained
    v = _frames(8, seed=800, H=384, W=384)
    frames = torch_dev.from_numpy(v.view(np.int16)).cuda()
    p = gt.make_params(9, characterize=True, **PRM)
    with gt.Context(p, v.shape[1], v.shape[2]) as ctx:
        plain = _rows(*ctx.detect_device(frames))
        ctx.set_learned(net, scope=scope)
        learned = _rows(*ctx.detect_device(frames))
        acc, cands = ctx.learned_stats(reset=True)
        ctx.set_learned(None)
        again = _rows(*ctx.detect_device(frames))
    for a, b in zip(plain, again):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    n_all = n_ok = 0
    for a, b in zip(plain, learned):
        bs = {int(r["seed"]): r for r in b}
        for r in a:
            n_all += 1
            q = bs.get(int(r["seed"]))
            if q is not None and abs(q["x"] - r["x"]) <= 1e-3 and abs(q["y"] - r["y"]) <= 1e-3:
                n_ok += 1
    frac = n_ok / n_all
    print(f"\n{scope}: accepted {acc}/{cands} proposals; {n_ok}/{n_all} particles within "
          "1e-3 px")
    assert acc > 0 and cands > 0
    assert frac >= 0.999, frac


def test_set_learned_rejects_other_paths(torch_dev, trained):
    import gpu_tracking_amd as gt
    net, _ = trained
    with gt.Context(gt.make_log_params(1.5, 4.0), 64, 64) as ctx:
        with pytest.raises(gt.GtError):
            ctx.set_learned(net)
    with gt.Context(gt.make_params(9, **PRM), 64, 64) as ctx:
        with pytest.raises(gt.GtError):
            ctx.set_learned(net, tol=float("nan"))
        with pytest.raises(ValueError):
            ctx.set_learned(net, scope="some")


@pytest.mark.parametrize("scope", ["moving", "all"])
def test_learned_path_on_the_full_workload_matches_the_oracle(torch_dev, oracle, trained, scope):
    # the evaluation shape and generator: 1024 x 1024 u16 frames of the full
    # workload (1000 particles), against the oracle's full walk
    import gpu_tracking_amd as gt
    from gpu_tracking_amd import synth
    net, _ = trained
    v = synth.video_numpy(synth.spec("full"), 2, 20261203 + 5000)
    print(f"[learned 1024] scope={scope} frames {v.shape}", flush=True)
    frames = torch_dev.from_numpy(v.view(np.int16)).cuda()
    p = gt.make_params(9, characterize=True, **PRM)
    with gt.Context(p, 1024, 1024) as ctx:
        ctx.set_learned(net, scope=scope)
        learned = _rows(*ctx.detect_device(frames))
        acc, cands = ctx.learned_stats(reset=True)
    n_all = n_ok = 0
    for t in range(v.shape[0]):
        op = oracle.derive_trackpy(1024, 1024, diameter=9, **PRM)
        ref, _ = oracle.detect_frame(v[t], op, characterize=True)
        bs = {int(r["seed"]): r for r in learned[t]}
        for r in ref:
            n_all += 1
            q = bs.get(int(r["seed"]))
            if q is not None and abs(q["x"] - r["x"]) <= 1e-3 and abs(q["y"] - r["y"]) <= 1e-3:
                n_ok += 1
    frac = n_ok / n_all
    print(f"\n1024^2 {scope}: accepted {acc}/{cands} proposals; {n_ok}/{n_all} oracle particles "
          "within 1e-3 px")
    assert acc > 0 and n_all > 1500
    assert frac >= 0.999, frac
