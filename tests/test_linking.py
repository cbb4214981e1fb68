"""Linking: the oracle's restatement of src/linking.rs pinned against brute
force and invariants, and the product's C++ linker (gt_link) against the
oracle, id for id.  All on the CPU."""
import itertools

import numpy as np
import pytest

from oracle import linker as OL


def _cost(src, dest, assign, r2):
    """Total cost of an assignment dest -> src index (or None)."""
    c = np.float64(0)
    for d, s in enumerate(assign):
        if s is None:
            c += r2
        else:
            c += float(OL._d2(dest[d], src[s][0]))
    return c


def _brute(src, dest, r):
    r2 = float(np.float32(np.float32(r) * np.float32(r)))
    best = None
    opts = [[None] + [s for s in range(len(src)) if OL._d2(dest[d], src[s][0]) < np.float32(r2)]
            for d in range(len(dest))]
    for assign in itertools.product(*opts):
        used = [s for s in assign if s is not None]
        if len(used) != len(set(used)):
            continue
        c = _cost(src, dest, assign, r2)
        if best is None or c < best:
            best = c
    return best


@pytest.mark.parametrize("seed", range(12))
def test_oracle_subnet_assignment_is_optimal(seed):
    rng = np.random.default_rng(seed)
    ns, nd = rng.integers(1, 5), rng.integers(1, 5)
    src = [((np.float32(rng.uniform(0, 4)), np.float32(rng.uniform(0, 4))), 100 + k)
           for k in range(ns)]
    dest = [(np.float32(rng.uniform(0, 4)), np.float32(rng.uniform(0, 4))) for _ in range(nd)]
    r = 2.5
    ids, unused, counter = OL.link(src, dest, None, r, 1000)
    idmap = {e[1]: k for k, e in enumerate(src)}
    assign = [idmap.get(i) for i in ids]
    r2 = float(np.float32(np.float32(r) * np.float32(r)))
    assert abs(_cost(src, dest, assign, r2) - _brute(src, dest, r)) < 1e-4
    # linked sources are distinct, new ids count up from the counter in dest order
    linked = [i for i in ids if i < 1000]
    assert len(linked) == len(set(linked))
    assert [i for i in ids if i >= 1000] == list(range(1000, counter))
    assert set(e[1] for e in unused) == set(e[1] for e in src) - set(linked)


def _tracks(rng, n, T, step, blink=0.0, box=200.0):
    pos = rng.uniform(10, box - 10, (n, 2))
    frames, truth = [], []
    for t in range(T):
        if t:
            pos = pos + rng.normal(0, step, pos.shape)
        vis = rng.random(n) >= blink
        frames.append(pos[vis].astype(np.float32))
        truth.append(np.flatnonzero(vis))
    return frames, truth


def test_oracle_keeps_ids_of_separated_particles():
    rng = np.random.default_rng(1)
    frames, truth = _tracks(rng, 30, 12, 0.3, box=400.0)
    lk = OL.Linker(3.0, 0)
    first = None
    for t, f in enumerate(frames):
        ids = lk.advance(f)
        if first is None:
            first = ids
        assert ids == first          # every particle keeps its id
    books = lk.finish()
    assert all(d == 12 for _, d in books)


def test_oracle_memory_and_durations():
    # particle A: frames 0-1, absent 2-3, back at 4; B: frames 0-5; C: 0-1 then gone
    A, B, C = (10.0, 10.0), (50.0, 50.0), (90.0, 90.0)
    seq = [[A, B, C], [A, B, C], [B], [B], [A, B], [B]]
    for memory, a_back in [(2, True), (1, False)]:
        lk = OL.Linker(5.0, memory)
        out = [lk.advance(f) for f in seq]
        a0 = out[0][0]
        if a_back:
            assert out[4][0] == a0       # within memory: same id
        else:
            assert out[4][0] != a0       # beyond memory: a new particle
        books = lk.finish()
        # duration = frames from first to last appearance, inclusive
        assert books[out[0][1]][1] == 6              # B
        assert books[out[0][2]][1] == 2              # C
        if a_back:
            assert books[a0][1] == 5                 # A: frames 0..4
        else:
            assert books[a0][1] == 2 and books[out[4][0]][1] == 1


def test_oracle_too_dense():
    src = [((np.float32(i * 0.01), np.float32(0)), i) for i in range(80)]
    dest = [(np.float32(i * 0.01), np.float32(0.001)) for i in range(80)]
    with pytest.raises(OL.TooDenseToLink):
        OL.link(src, dest, None, 5.0, 80)


# --------------------------------------------------------------------------
# product (gt_link, C++) against the oracle
def _as_rows(frames):
    rows = []
    for t, f in enumerate(frames):
        for p in f:
            rows.append((t, p[0], p[1]))
    return np.array(rows, np.float32).reshape(-1, 3)


@pytest.mark.parametrize("seed,n,T,step,blink,sr,mem", [
    (0, 60, 15, 1.0, 0.05, 4.0, 0),
    (1, 80, 20, 1.5, 0.1, 5.0, 2),
    (2, 40, 10, 2.5, 0.0, 6.0, 1),    # crowded: subnetworks of several particles
    (3, 120, 8, 0.8, 0.2, 3.0, 3),
])
def test_gt_link_matches_oracle(seed, n, T, step, blink, sr, mem):
    from gpu_tracking_amd import _native as N
    rng = np.random.default_rng(seed)
    frames, _ = _tracks(rng, n, T, step, blink, box=120.0)
    rows = _as_rows(frames)
    ref = OL.link_frames(rows, sr, mem)
    ids = np.zeros(len(rows), np.uint64)
    N.check(N.lib().gt_link(rows.ctypes.data, len(rows), sr, mem, ids.ctypes.data))
    assert ids.tolist() == ref


@pytest.mark.parametrize("seed,box,shift,sr", [
    (10, 2.0e5, 0.0, 3.0),       # strips far more than points: the sorted-strip search
    (11, 120.0, -5000.0, 4.0),   # negative rows and columns
    (12, 120.0, 0.0, 0.05),      # a radius much smaller than the spacing
    (13, 60.0, 0.0, 1.0e4),      # a radius spanning the field: one strip
])
def test_gt_link_matches_oracle_strip_layouts(seed, box, shift, sr):
    from gpu_tracking_amd import _native as N
    rng = np.random.default_rng(seed)
    n = 6 if sr > 100 else 40
    frames, _ = _tracks(rng, n, 8, 1.0, 0.1, box=box)
    rows = _as_rows(frames)
    rows[:, 1:] += np.float32(shift)
    ref = OL.link_frames(rows, sr, 1)
    ids = np.zeros(len(rows), np.uint64)
    N.check(N.lib().gt_link(rows.ctypes.data, len(rows), sr, 1, ids.ctypes.data))
    assert ids.tolist() == ref


def test_gt_link_groups_and_errors():
    import gpu_tracking_amd as gt
    import pandas as pd
    from gpu_tracking_amd import _native as N
    # frames 2 and 5 only: an empty frame 0 comes first, skipped numbers are not frames
    df = pd.DataFrame({"frame": [2, 2, 5, 5], "y": [1.0, 20.0, 1.5, 20.5], "x": [1, 1, 1, 1]})
    out = gt.link(df, 2.0, 0)
    assert list(out["particle"]) == [0, 1, 0, 1]
    ref = OL.link_frames(df[["frame", "y", "x"]].to_numpy(), 2.0, 0)
    assert list(out["particle"]) == ref
    with pytest.raises(N.GtError) as e:
        gt.link(np.array([[1, 0, 0], [0, 1, 1]], np.float32), 2.0)
    assert e.value.status == 6
    dense = np.array([[0, i * 0.01, 0] for i in range(80)] + [[1, i * 0.01, 0.001]
                                                               for i in range(80)], np.float32)
    with pytest.raises(N.GtError) as e:
        gt.link(dense, 5.0)
    assert e.value.status == 18


def _connect_sets(rng, nframes, p_frame, n, box=30.0):
    """Two point sets over frames 0..nframes-1 (each frame present in a set
    with probability p_frame), set 2 a jittered, thinned copy of set 1 plus
    extra points where the frame is in both."""
    s1, s2 = [], []
    for f in range(nframes):
        pts = rng.uniform(0, box, (rng.integers(0, n + 1), 2)).astype(np.float32)
        in1, in2 = rng.random() < p_frame, rng.random() < p_frame
        if in1:
            s1 += [(f, y, x) for y, x in pts]
        if in2:
            keep = pts[rng.random(len(pts)) < 0.8] + rng.normal(0, 0.6, (1, 2)).astype(np.float32)
            extra = rng.uniform(0, box, (rng.integers(0, 3), 2))
            s2 += [(f, y, x) for y, x in np.concatenate([keep, extra])]
    return (np.array(s1, np.float32).reshape(-1, 3), np.array(s2, np.float32).reshape(-1, 3))


def test_oracle_connect_literal_double_fill():
    # set 2 skips frame 2 and overtakes it with frame 3: the literal
    # sync_iterators gives frame 3's point two ids (4 ids for 3 rows)
    a = np.array([[0, 1, 1], [2, 5, 5]], np.float32)
    b = np.array([[0, 1.2, 1], [1, 3, 3], [3, 7, 7]], np.float32)
    i1, i2 = OL.connect_frames(a, b, 2.0, fixed=False)
    assert len(i1) == 2 and len(i2) == 4
    assert OL.connect_frames(a, b, 2.0) == ([0, 2], [0, 1, 3])


@pytest.mark.parametrize("seed", range(16))
def test_oracle_connect_pinned(seed):
    rng = np.random.default_rng(100 + seed)
    r = 2.0
    s1, s2 = _connect_sets(rng, int(rng.integers(1, 7)), 0.7, 4)
    i1, i2 = OL.connect_frames(s1, s2, r)
    lit1, lit2 = OL.connect_frames(s1, s2, r, fixed=False)
    assert len(i1) == len(s1) and len(i2) == len(s2)
    if len(lit1) == len(s1) and len(lit2) == len(s2):
        # where the literal restatement is well formed the two agree
        assert (lit1, lit2) == (i1, i2)
    assert len(set(i1)) == len(i1) and len(set(i2)) == len(i2)
    # every id is handed out once, counting up with no gaps
    assert sorted(set(i1) | set(i2)) == list(range(len(set(i1) | set(i2))))
    r2 = float(np.float32(np.float32(r) * np.float32(r)))
    pos1 = {i: row for i, row in zip(i1, s1)}
    for f in sorted(set(s2[:, 0].astype(int))):
        d = [tuple(row[1:]) for row in s2 if int(row[0]) == f]
        ids = [i for i, row in zip(i2, s2) if int(row[0]) == f]
        src_rows = [row for row in s1 if int(row[0]) == f]
        src = [((row[1], row[2]), i) for i, row in zip(i1, s1) if int(row[0]) == f]
        for i, q in zip(ids, d):
            if i in pos1:   # linked: same frame, within the search range
                p = pos1[i]
                assert int(p[0]) == f and OL._d2(q, (p[1], p[2])) < np.float32(r2)
        if src_rows and d and len(src) <= 4 and len(d) <= 4:
            # set 1 is the source side, set 2 the destinations: the linked
            # pairs are a minimum-cost assignment
            idmap = {e[1]: k for k, e in enumerate(src)}
            assign = [idmap.get(i) for i in ids]
            assert abs(_cost(src, d, assign, r2) - _brute(src, d, r)) < 1e-4


@pytest.mark.parametrize("seed", range(10))
def test_gt_connect_matches_oracle(seed):
    from gpu_tracking_amd import _native as N
    rng = np.random.default_rng(300 + seed)
    s1, s2 = _connect_sets(rng, int(rng.integers(1, 12)), [1.0, 0.8, 0.5][seed % 3], 12)
    if seed == 9:
        s2 = s2[:0]
    ref1, ref2 = OL.connect_frames(s1, s2, 2.5)
    ids1 = np.zeros(len(s1), np.uint64)
    ids2 = np.zeros(len(s2), np.uint64)
    N.check(N.lib().gt_connect(s1.ctypes.data, len(s1), s2.ctypes.data, len(s2), 2.5,
                               ids1.ctypes.data, ids2.ctypes.data))
    assert ids1.tolist() == ref1 and ids2.tolist() == ref2


def test_connect_wrapper():
    import gpu_tracking_amd as gt
    import pandas as pd
    from gpu_tracking_amd import _native as N
    a = pd.DataFrame({"frame": [0, 0, 2], "y": [1.0, 10.0, 5.0], "x": [1.0, 1.0, 5.0],
                      "mass": [5.0, 6.0, 7.0]})
    b = pd.DataFrame({"frame": [0, 1, 2, 2], "y": [1.3, 3.0, 5.2, 20.0],
                      "x": [1.0, 3.0, 5.0, 1.0], "mass": [1.0, 2.0, 3.0, 4.0]})
    o1, o2 = gt.connect(a, b, 1.0, merge=False)
    assert list(o1["connect particle"]) == [0, 1, 3]
    assert list(o2["connect particle"]) == [0, 2, 3, 4]
    m = gt.connect(a, b, 1.0)
    assert len(m) == 5 and {"frame_x", "frame_y", "mass_x", "mass_y"} <= set(m.columns)
    both = m.dropna()
    assert sorted(both["mass_x"]) == [5.0, 7.0] and sorted(both["mass_y"]) == [1.0, 3.0]
    i1, i2 = gt.connect(a[["frame", "y", "x"]].to_numpy(), b, 1.0)
    assert i1.tolist() == [0, 1, 3] and list(i2["connect particle"]) == [0, 2, 3, 4]
    with pytest.raises(N.GtError) as e:
        gt.connect(np.array([[1, 0, 0], [0, 1, 1]], np.float32), b, 1.0)
    assert e.value.status == 6


def test_unique_particle_ids():
    import gpu_tracking_amd as gt
    import pandas as pd
    a = pd.DataFrame({"particle": [0, 1, 1]})
    b = pd.DataFrame({"particle": [0, 1, 1]})
    gt.unique_particle_ids(a)
    gt.unique_particle_ids(b)
    assert a["particle"][1] == a["particle"][2] and a["particle"][0] != a["particle"][1]
    assert not set(a["particle"]) & set(b["particle"])
    assert all(v.startswith(str(i)) for v, i in zip(a["particle"], [0, 1, 1]))
