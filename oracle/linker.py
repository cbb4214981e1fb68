"""CPU restatement of gpu_tracking's frame-to-frame linker -- TEST
INFRASTRUCTURE ONLY (the product's linker is gpu_tracking_amd/csrc/gt_linker.cpp).

Follows src/linking.rs: Linker::advance / finish (:435-521), link (:564-816)
and linker_all (:818-831), with within_radius_rd2's strict d^2 < r^2 test
(:208-247).  Pure Python, meant for the small cases of the tests.

The one free choice: the reference lists a source's candidate destinations
in the order its kd-tree returns them, which sets the depth-first order of
the subnetworks and so which of several equal-cost assignments wins.  Here
they are listed in ascending destination index (the product does the same).
Score sums are f32, in the reference's order of additions and subtractions.
"""
from __future__ import annotations

from collections import deque

import numpy as np

f32 = np.float32


class TooDenseToLink(Exception):
    pass


def _d2(item, query):
    """within_radius_rd2's distance: (0 + dx^2) + dy^2 in f32, item - query."""
    dx = f32(item[0]) - f32(query[0])
    dy = f32(item[1]) - f32(query[1])
    d = f32(0.0)
    d = f32(d + f32(dx * dx))
    d = f32(d + f32(dy * dy))
    return d


def link(src, dest, memory_start_idx, radius, counter):
    """link (linking.rs:564-816).  src: list of ((x, y), id); dest: list of
    (x, y).  Returns (ids for dest, unused sources, new counter)."""
    radius = f32(radius)
    r2 = f32(radius * radius)
    ns, nd = len(src), len(dest)
    s2d = [[] for _ in range(ns)]
    d2s = [[] for _ in range(nd)]
    for s, (pt, _) in enumerate(src):
        for d in range(nd):
            dd = _d2(dest[d], pt)
            if dd < r2:
                s2d[s].append((d, dd))
                d2s[d].append((s, dd))
    visited = [[False] * ns, [False] * nd]

    def recurse(node, kind, path):
        if visited[kind][node]:
            return
        visited[kind][node] = True
        path[kind].append(node)
        if kind == 0:
            for d, _ in s2d[node]:
                recurse(d, 1, path)
        else:
            for s, _ in d2s[node]:
                recurse(s, 0, path)

    paths = []
    load = 0
    output = [None] * nd
    for i in range(nd):
        path = [[], []]
        recurse(i, 1, path)
        if len(path[1]) == 1:
            if len(path[0]) == 1:
                output[i] = src[path[0][0]][1]
            elif len(path[0]) > 1:
                paths.append(path)
        elif len(path[1]) > 0:
            load += len(path[0]) * len(path[1])
            paths.append(path)
    if load > 5000:
        raise TooDenseToLink()

    def recurse2(progress, path, st):
        cur = path[1][progress]
        cands = [(src[s][1], dd) for s, dd in d2s[cur]] + [(None, r2)]
        for ident, cost in cands:
            if st["score"] > st["best"]:
                return
            st["score"] = f32(st["score"] + cost)
            if ident is not None:
                if ident in st["used"]:
                    st["score"] = f32(st["score"] - cost)
                    continue
                st["used"][ident] = cur
            else:
                st["nulls"].add(cur)
            if progress == len(path[1]) - 1:
                if st["score"] < st["best"]:
                    st["best"] = st["score"]
                    for k, v in st["used"].items():
                        output[v] = k
                    for v in st["nulls"]:
                        output[v] = None
            else:
                recurse2(progress + 1, path, st)
            if ident is not None:
                del st["used"][ident]
            else:
                st["nulls"].discard(cur)
            st["score"] = f32(st["score"] - cost)

    for path in paths:
        recurse2(0, path, {"score": f32(0.0), "best": f32(np.inf), "used": {}, "nulls": set()})
    ids = []
    for v in output:
        if v is None:
            ids.append(counter)
            counter += 1
        else:
            ids.append(v)
    used = set(ids)
    lim = memory_start_idx if memory_start_idx is not None else len(src)
    unused = [e for e in src[:lim] if e[1] not in used]
    return ids, unused, counter


class Linker:
    """Linker (linking.rs:346-521)."""

    def __init__(self, search_range, memory):
        self.search_range = f32(search_range)
        self.memory = int(memory)
        self.prev = []
        self.prev_n = 0
        self.memvec = deque([[] for _ in range(self.memory)]) if self.memory > 0 else None
        self.frame_idx = 0
        self.part_idx = 0
        self.books = []   # [start, duration] per particle id

    def advance(self, frame):
        frame = [(f32(x), f32(y)) for x, y in frame]
        n = len(frame)
        mstart = None
        if self.memvec is not None:
            seen = set(e[1] for e in self.prev)
            for vec in self.memvec:
                for e in vec:
                    if e[1] not in seen:
                        self.prev.append(e)
                        seen.add(e[1])
            mstart = self.prev_n
        before = self.part_idx
        ids, unused, self.part_idx = link(self.prev, frame, mstart, self.search_range,
                                          self.part_idx)
        for _ in range(self.part_idx - before):
            self.books.append([self.frame_idx, 0])
        self.prev = [(p, i) for p, i in zip(frame, ids)]
        if self.memvec is not None:
            self.memvec.append(unused)
            for e in self.memvec.popleft():
                b = self.books[e[1]]
                b[1] = (self.frame_idx - b[0]) - len(self.memvec)
        else:
            for e in unused:
                b = self.books[e[1]]
                b[1] = self.frame_idx - b[0]
        self.prev_n = n
        self.frame_idx += 1
        return ids

    def finish(self):
        if self.memvec is not None:
            L = len(self.memvec)
            for i, vec in enumerate(self.memvec):
                for e in vec:
                    b = self.books[e[1]]
                    b[1] = (self.frame_idx - b[0]) - (L - i)
        for e in self.prev:
            b = self.books[e[1]]
            b[1] = self.frame_idx - b[0]
        return [tuple(b) for b in self.books]


def link_frames(points, search_range, memory):
    """linker_all over a (n, 3) float array [frame, y, x] sorted by frame
    (FrameSubsetter grouping: consecutive groups of equal frame value, the
    first group being frame 0 even when empty).  Returns one id per row."""
    pts = np.asarray(points, np.float32)
    lk = Linker(search_range, memory)
    out = []
    cur = f32(0.0)
    group = []
    for row in pts:
        fr = row[0]
        if fr > cur:
            out.extend(lk.advance(group))
            group = []
            cur = fr
        if fr < cur:
            raise ValueError("not sorted by frame")
        group.append((row[1], row[2]))
    if group:
        out.extend(lk.advance(group))
    return out


def frame_groups(points):
    """FrameSubsetter's linking groups (linking.rs:94-155) of a (n, 3) float
    array [frame, y, x]: consecutive rows of one frame value form a group,
    keyed by the frame value cast to an unsigned integer; when the first row's
    frame is above 0 an empty group keyed 0 comes first; no rows, no groups."""
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    groups = []
    cur = f32(0.0)
    group = []
    for row in pts:
        fr = row[0]
        if fr > cur:
            groups.append((int(cur), group))
            group = []
            cur = fr
        if fr < cur:
            raise ValueError("not sorted by frame")
        group.append((row[1], row[2]))
    if group:
        groups.append((int(cur), group))
    return groups


def connect_frames(points1, points2, search_range, fixed=True):
    """connect_all (linking.rs:960-1012) with Linker::connect (:405-432),
    fill_base and sync_iterators (:833-958): frames present in both sets are
    linked (set 1's points take fresh ids, set 2's points the id of the set-1
    point they link to or a fresh one); a frame present in one set only takes
    fresh ids.  Ids count up in the order the frames are visited.

    fixed=False follows sync_iterators literally: when the lagging set skips
    past the frame it was catching up to (Ordering::Less, :945-955), the
    overtaking frame's points get ids there and again when that frame is
    later linked or filled, so that set's id list comes out longer than its
    rows.  fixed=True (the product's behaviour) gives the skipped frame its
    ids and carries the overtaking frame on, unfilled.
    Returns (ids1, ids2) as lists."""
    it = [iter(frame_groups(points1)), iter(frame_groups(points2))]
    res = [[], []]
    counter = [0]
    r = f32(search_range)

    def fill(side, item):
        for _ in item[1]:
            res[side].append(counter[0])
            counter[0] += 1

    def connect(a, b):
        src = []
        for p in a[1]:
            src.append(((f32(p[0]), f32(p[1])), counter[0]))
            counter[0] += 1
        dest = [(f32(p[0]), f32(p[1])) for p in b[1]]
        ids, _, counter[0] = link(src, dest, None, r, counter[0])
        res[0].extend(e[1] for e in src)
        res[1].extend(ids)

    def sync(hit_side, hit):
        # hit: the frame ahead, on side hit_side; the other side catches up
        while True:
            nxt = next(it[1 - hit_side], None)
            if nxt is None:
                fill(hit_side, hit)
                return None
            if hit[0] > nxt[0]:
                fill(1 - hit_side, nxt)
                continue
            if hit[0] == nxt[0]:
                return (hit, nxt) if hit_side == 0 else (nxt, hit)
            if not fixed:
                fill(1 - hit_side, nxt)
            fill(hit_side, hit)
            hit_side, hit = 1 - hit_side, nxt

    while True:
        a = next(it[0], None)
        b = next(it[1], None)
        if a is None and b is None:
            break
        if b is None:
            fill(0, a)
            continue
        if a is None:
            fill(1, b)
            continue
        if a[0] == b[0]:
            pair = (a, b)
        elif a[0] > b[0]:
            fill(1, b)
            pair = sync(0, a)
        else:
            fill(0, a)
            pair = sync(1, b)
        if pair is not None:
            connect(*pair)
    return res[0], res[1]
