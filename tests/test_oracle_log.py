"""Pins the LoG part of the CPU oracle against mathematics it must obey.

The reference's LoG path (separable_log.wgsl, logspace_max.wgsl,
prune_blobs_log in execute_gpu.rs) has no tests or fixtures of its own; each
oracle function is checked against an independent statement: the analytic
scale-normalised LoG kernel, a direct 2-D sum in f64, the scale at which a
Gaussian blob's response peaks, a brute-force scale-space maximum, the
circle-intersection area formula, and the defining property of the greedy
blob pruning.
"""
import math

import numpy as np
import pytest

SQRT2 = np.float32(np.sqrt(np.float32(2.0)))


# --------------------------------------------------------------------------
# radius scan (gpu_setup.rs:958-984)
@pytest.mark.parametrize("lo,hi,n", [(1.5, 6.0, 10), (1.0, 3.0, 5), (2.0, 2.5, 2), (3.0, 3.0, 4)])
def test_radii_match_linspace_and_logspace(oracle, lo, hi, n):
    lin = oracle.log_radii(lo, hi, n, False)
    assert lin.dtype == np.float32 and len(lin) == n
    np.testing.assert_allclose(lin, np.linspace(lo, hi, n), rtol=2e-6, atol=2e-6)
    lg = oracle.log_radii(lo, hi, n, True)
    np.testing.assert_allclose(lg, np.logspace(np.log10(lo), np.log10(hi), n), rtol=4e-6)
    # ascending (or constant) and starting where asked
    assert np.all(np.diff(lin) >= 0) and np.all(np.diff(lg) >= 0)
    assert lin[0] == np.float32(lo)


def test_radii_single_scale(oracle):
    assert list(oracle.log_radii(2.0, 5.0, 1, False)) == [np.float32(2.0)]
    r = oracle.log_radii(2.0, 5.0, 1, True)
    assert abs(r[0] - 2.0) < 1e-6


# --------------------------------------------------------------------------
# taps and norms (separable_log.wgsl:22-89)
@pytest.mark.parametrize("sigma", [0.7, 1.0607, 2.5, 4.2426])
def test_taps_closed_form(oracle, sigma):
    R, g, d, n1, n3 = oracle.log_taps(sigma)
    s = np.float64(np.float32(sigma))
    assert R == int(np.floor(4.0 * np.float32(sigma) + 0.5))
    x = np.arange(-R, R + 1, dtype=np.float64)
    e = np.exp(-x * x / (2 * s * s))
    # the exponent is an f32 quotient: relative error up to ~|x^2/2s^2| ulp
    np.testing.assert_allclose(g, e, rtol=2e-6, atol=1e-30)
    np.testing.assert_allclose(d, e * (x * x - s * s), rtol=2e-6, atol=1e-6)
    assert np.isclose(n1, 1 / (np.sqrt(2 * np.pi) * s), rtol=3e-7)
    assert np.isclose(n3, 1 / (np.sqrt(2 * np.pi) * s ** 3), rtol=5e-7)
    # symmetric taps
    assert np.array_equal(g, g[::-1]) and np.array_equal(d, d[::-1])


def _log_kernel(sigma, R):
    """Analytic scale-normalised LoG kernel -sigma^2 (G_xx + G_yy) sampled on
    [-R, R]^2, with G the unnormalised-sum 1-D Gaussian exp(-x^2/2s^2)/(sqrt(2pi) s)
    and sigma^2 G'' = G (x^2 - s^2) / s^2."""
    s = np.float64(sigma)
    x = np.arange(-R, R + 1, dtype=np.float64)
    G = np.exp(-x * x / (2 * s * s)) / (np.sqrt(2 * np.pi) * s)
    D = G * (x * x - s * s) / (s * s)
    return -(np.outer(D, G) + np.outer(G, D))


@pytest.mark.parametrize("sigma", [1.0607, 2.0, 3.3])
def test_log_plane_impulse_response(oracle, sigma):
    sigma = float(np.float32(sigma))
    R = oracle.log_taps(sigma)[0]
    H = W = 2 * R + 9
    f = np.zeros((H, W), np.float32)
    f[H // 2, W // 2] = 1.0
    L = oracle.log_plane(f, sigma)
    K = _log_kernel(sigma, R)
    got = L[H // 2 - R:H // 2 + R + 1, W // 2 - R:W // 2 + R + 1].astype(np.float64)
    np.testing.assert_allclose(got, K, rtol=2e-5, atol=2e-7 * np.abs(K).max())
    # nothing outside the kernel's support
    mask = np.ones_like(L, bool)
    mask[H // 2 - R:H // 2 + R + 1, W // 2 - R:W // 2 + R + 1] = False
    assert np.all(L[mask] == 0)


@pytest.mark.parametrize("sigma,shape", [(1.0607, (23, 19)), (2.2, (31, 40))])
def test_log_plane_direct_2d_sum(oracle, sigma, shape):
    # separable passes against a direct zero-padded 2-D sum in f64
    sigma = float(np.float32(sigma))
    rng = np.random.default_rng(7)
    f = rng.uniform(0, 100, shape).astype(np.float32)
    R = oracle.log_taps(sigma)[0]
    K = _log_kernel(sigma, R)
    H, W = shape
    fp = np.zeros((H + 2 * R, W + 2 * R))
    fp[R:R + H, R:R + W] = f
    ref = np.zeros((H, W))
    for i in range(2 * R + 1):
        for j in range(2 * R + 1):
            ref += K[i, j] * fp[i:i + H, j:j + W]
    L = oracle.log_plane(f, sigma)
    np.testing.assert_allclose(L, ref, rtol=0, atol=2e-5 * np.abs(f).max())


def test_log_scale_selection_peaks_at_blob_sigma(oracle):
    # -s^2 lap(G_s * blob_b) at the blob centre ~ s^2 / (s^2 + b^2)^2: largest at s = b
    b = 2.5
    n = 61
    y, x = np.mgrid[0:n, 0:n] - n // 2
    blob = np.exp(-(x * x + y * y) / (2 * b * b)).astype(np.float32) * 1000
    sig = np.linspace(1.6, 3.6, 21)
    resp = [oracle.log_plane(blob, float(s))[n // 2, n // 2] for s in sig]
    best = sig[int(np.argmax(resp))]
    assert abs(best - b) <= 0.1 + 1e-9
    assert min(resp) > 0


# --------------------------------------------------------------------------
# scale-space maxima (logspace_max.wgsl:62-154)
def _brute_max(B, M, T, P, thr):
    H, W = M.shape
    out = []
    for i in range(H):
        for j in range(W):
            if P[i, j] < thr:
                continue
            c = M[i, j]
            if c <= 0:
                continue
            ok = True
            for pl in (M, B, T):
                if pl is None:
                    continue
                for a in (-1, 0, 1):
                    for b in (-1, 0, 1):
                        if a == 0 and b == 0:
                            continue
                        x, y = i + a, j + b
                        if not (0 <= x < H and 0 <= y < W) or pl[x, y] >= c:
                            ok = False
            if ok:
                out.append(i * W + j)
    return np.array(out, np.uint32)


@pytest.mark.parametrize("which", ["all", "first", "last", "single"])
def test_log_max_brute_force(oracle, which):
    rng = np.random.default_rng(11)
    H, W = 17, 23
    planes = [np.round(rng.normal(0, 3, (H, W))).astype(np.float32) for _ in range(3)]
    # plateaus and ties in the same plane and across planes
    planes[1][5, 5] = planes[1][5, 6] = 9.0
    planes[1][10, 10] = 12.0
    planes[0][10, 10] = 12.0   # directly below: does not count
    planes[2][9, 10] = 12.0    # a neighbour above: does
    planes[1][3, 15] = 20.0
    P = rng.uniform(0, 1, (H, W)).astype(np.float32)
    B, M, T = planes
    if which == "first":
        B = None
    elif which == "last":
        T = None
    elif which == "single":
        B = T = None
    got = oracle.log_max(B, M, T, P)
    ref = _brute_max(B, M, T, P, 0.0)
    assert np.array_equal(got, ref)
    if T is not None:
        assert 10 * W + 10 not in set(got.tolist())   # the top plane's neighbour ties it
    else:
        assert 3 * W + 15 in set(got.tolist())


def test_log_max_rejects_border_and_nonpositive(oracle):
    H, W = 9, 9
    M = np.zeros((H, W), np.float32)
    M[0, 4] = 5.0      # on the border: a neighbour is outside the frame
    M[4, 4] = -1.0     # a negative dip is never a maximum
    M[2, 2] = 3.0
    P = np.ones((H, W), np.float32)
    got = oracle.log_max(None, M, None, P).tolist()
    assert got == [2 * W + 2]
    # processed below the snr threshold rejects
    P[2, 2] = 0.25
    assert oracle.log_max(None, M, None, P, 0.5).tolist() == []


# --------------------------------------------------------------------------
# blob overlap (execute_gpu.rs:597-630) and pruning (:592-712)
def _lens_area(d, r1, r2):
    a1 = r1 * r1 * math.acos((d * d + r1 * r1 - r2 * r2) / (2 * d * r1))
    a2 = r2 * r2 * math.acos((d * d + r2 * r2 - r1 * r1) / (2 * d * r2))
    k = 0.5 * math.sqrt((-d + r1 + r2) * (d + r1 - r2) * (d - r1 + r2) * (d + r1 + r2))
    return a1 + a2 - k


@pytest.mark.parametrize("r1,r2,d", [(1.0, 1.0, 1.0), (2.0, 1.5, 2.2), (3.0, 1.0, 2.5),
                                     (2.5, 2.5, 4.9)])
def test_blob_overlap_lens_area(oracle, r1, r2, d):
    ref = _lens_area(d, r1, r2) / (math.pi * min(r1, r2) ** 2)
    assert abs(oracle.blob_overlap(r1, r2, d) - ref) < 2e-6
    assert abs(oracle.blob_overlap(r2, r1, d) - ref) < 2e-6


def test_blob_overlap_limits(oracle):
    assert oracle.blob_overlap(1.0, 1.0, 2.5) == 0.0
    assert oracle.blob_overlap(3.0, 1.0, 1.5) == 1.0      # the small disk lies inside
    # two equal unit disks at distance 1: (2 acos(1/2) - sqrt(3)/2) / pi
    want = (2 * math.acos(0.5) - math.sqrt(3) / 2) / math.pi
    assert abs(oracle.blob_overlap(1.0, 1.0, 1.0) - want) < 1e-6


def _rows(oracle, pts):
    r = np.zeros(len(pts), oracle.ROW_DTYPE)
    for k, (x, y, rad, m) in enumerate(pts):
        r[k]["x"], r[k]["y"], r[k]["r"], r[k]["mass"] = x, y, rad, m
        r[k]["seed"] = k
    return r


def test_prune_order_and_greedy_property(oracle):
    rng = np.random.default_rng(3)
    n = 300
    pts = [(rng.uniform(0, 60), rng.uniform(0, 60), float(rng.choice([1.5, 2.0, 3.0])),
            float(rng.integers(1, 6))) for _ in range(n)]
    rows = _rows(oracle, pts)
    for thr in (0.0, 0.3):
        kept = oracle.prune_blobs(rows, thr)
        # survivors come out sorted by r desc then mass desc (stable in seed)
        key = [(-k["r"], -k["mass"], k["seed"]) for k in kept]
        assert key == sorted(key)
        # greedy in priority order: no two survivors overlap above thr, and
        # every dropped blob overlaps a survivor ranked before it above thr
        order = sorted(range(n), key=lambda k: (-rows[k]["r"], -rows[k]["mass"], k))
        rank = {k: i for i, k in enumerate(order)}
        ks = set(int(s) for s in kept["seed"])

        def ov(a, b):
            d = math.hypot(rows[a]["x"] - rows[b]["x"], rows[a]["y"] - rows[b]["y"])
            if d >= 2 * max(rows[a]["r"], rows[b]["r"]):
                return 0.0
            return oracle.blob_overlap(float(rows[a]["r"]), float(rows[b]["r"]), d)

        kl = sorted(ks)
        for i, a in enumerate(kl):
            for b in kl[i + 1:]:
                assert ov(a, b) <= thr + 1e-6
        for k in set(range(n)) - ks:
            assert any(ov(s, k) > thr - 1e-6 and rank[s] < rank[k] for s in ks), k


def test_prune_threshold_one_keeps_all(oracle):
    rows = _rows(oracle, [(0, 0, 2, 1), (0.5, 0, 2, 2)])
    assert len(oracle.prune_blobs(rows, 1.0)) == 2
    assert len(oracle.prune_blobs(rows, 0.0)) == 1
    assert oracle.prune_blobs(rows, 0.0)[0]["mass"] == 2     # the heavier of equal radii


# --------------------------------------------------------------------------
# whole frame
def test_log_frame_finds_blobs_at_their_scale(oracle):
    from gpu_tracking_amd import synth
    s = synth.spec("log", nrows=96, ncols=96, n_particles=6, sigma=(2.0, 2.0),
                   amp=("uniform", 1500.0, 1500.0), noise="gauss:0.5", min_dist=25.0, edge=15.0)
    v = synth.video_numpy(s, 1, seed=4)
    traj, *_ = synth.trajectories(s, 1, 4)
    p = oracle.derive_log(96, 96, 6.0)
    radii = oracle.log_radii(1.5, 6.0, 10, False)
    rows, nc = oracle.log_detect_frame(v[0], p, radii, overlap_threshold=0.0)
    # each particle is found within a pixel, at a radius near sigma * sqrt(2)
    pos = traj[0]
    for y, x in pos:
        d = np.hypot(rows["x"] - y, rows["y"] - x)
        k = int(np.nanargmin(d))     # zero-mass walks end at NaN
        assert d[k] < 1.0
        assert abs(rows["r"][k] - 2.0 * math.sqrt(2)) <= 1.0
    assert np.all(rows["r"] > 0) and nc >= len(pos)


# --------------------------------------------------------------------------
# threshold_rel (no reference counterpart: skimage blob_log's relative
# threshold; the reference allocates a global_max it never writes,
# gpu_setup.rs:481-488)
@pytest.mark.parametrize("rel", [0.05, 0.3])
def test_threshold_rel_brute_force(oracle, rel):
    from gpu_tracking_amd import synth
    H, W = 40, 44
    s = synth.spec("log", nrows=H, ncols=W, n_particles=8, sigma=(1.0, 3.0),
                   amp=("uniform", 200.0, 2000.0), noise="gauss:5", min_dist=8.0, edge=6.0)
    frame = synth.video_numpy(s, 1, seed=9)[0]
    radii = oracle.log_radii(1.5, 6.0, 5, True)
    planes = [oracle.log_plane(frame, float(r) / math.sqrt(2.0)) for r in np.float32(radii)]
    gmax = max(0.0, max(float(pl.max()) for pl in planes))
    thr = np.float32(rel) * np.float32(gmax)
    P = oracle.bandpass(frame, oracle.derive_log(H, W, 6.0))
    want, unfiltered = set(), set()
    for m in range(len(radii)):
        B = planes[m - 1] if m > 0 else None
        T = planes[m + 1] if m + 1 < len(radii) else None
        for c in _brute_max(B, planes[m], T, P, 0.0).tolist():
            unfiltered.add(m * H * W + c)
            if planes[m].flat[c] > thr:
                want.add(m * H * W + c)
    # minmass 0 keeps every walk and overlap 1 prunes nothing, so the rows'
    # seeds are the candidate set
    p = oracle.derive_log(H, W, 6.0, threshold_rel=rel)
    rows, nc = oracle.log_detect_frame(frame, p, radii, characterize=False, overlap_threshold=1.0)
    assert nc == len(want) and set(rows["seed"].tolist()) == want
    assert len(want) < len(unfiltered)          # the rule removed candidates
    off, nc_off = oracle.log_detect_frame(frame, oracle.derive_log(H, W, 6.0), radii,
                                          characterize=False, overlap_threshold=1.0)
    assert nc_off == len(unfiltered) and set(off["seed"].tolist()) == unfiltered
    # the kept candidates walk exactly as without the rule
    keep = np.isin(off["seed"], rows["seed"])
    assert np.array_equal(np.sort(off[keep], order="seed").view(np.uint32),
                          np.sort(rows, order="seed").view(np.uint32))
