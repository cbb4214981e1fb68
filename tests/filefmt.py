"""Byte-level writers for the video-file tests: TIFF pages in layouts
Pillow does not produce (tiles, BigTIFF, big-endian, signed samples,
hand-encoded PackBits) and the ETS container as MinimalETSParser reads it
(src/decoderiter.rs:351-408)."""
import struct

import numpy as np


def tiff_bytes(pages, *, big=False, be=False, tile=None, fmt=1, rps=None, packbits=False):
    """pages: list of 2-D arrays of one dtype.  Writes strips (rps rows each)
    or tiles (tile = (th, tw), edge tiles padded), optionally PackBits."""
    e = ">" if be else "<"
    a0 = pages[0]
    H, W = a0.shape
    bits = a0.dtype.itemsize * 8
    out = bytearray()
    out += (b"MM" if be else b"II")
    if big:
        out += struct.pack(e + "HHHQ", 43, 8, 0, 0)
    else:
        out += struct.pack(e + "HI", 42, 0)
    first_ptr = 8 if big else 4
    prev_next_ptr = first_ptr

    def enc(b):
        if not packbits:
            return b
        # literal runs of at most 128 bytes, and a repeat run when 3+ equal
        res = bytearray()
        i = 0
        while i < len(b):
            j = i
            while j < len(b) and j - i < 128 and b[j] == b[i]:
                j += 1
            if j - i >= 3:
                res += struct.pack("b", -(j - i - 1)) + b[i:i + 1]
                i = j
                continue
            k = i
            while k < len(b) and k - i < 128 and not (k + 2 < len(b) and b[k] == b[k + 1] == b[k + 2]):
                k += 1
            k = max(k, i + 1)
            res += struct.pack("B", k - i - 1) + b[i:k]
            i = k
        return bytes(res)

    for a in pages:
        data = a.astype(a.dtype.newbyteorder(e))
        segs = []
        if tile:
            th, tw = tile
            for ty in range(0, H, th):
                for tx in range(0, W, tw):
                    t = np.zeros((th, tw), data.dtype)
                    blk = data[ty:ty + th, tx:tx + tw]
                    t[:blk.shape[0], :blk.shape[1]] = blk
                    segs.append(enc(t.tobytes()))
        else:
            r = rps or H
            for y in range(0, H, r):
                segs.append(enc(data[y:y + r].tobytes()))
        offs = []
        for sgm in segs:
            if len(out) % 2:
                out += b"\0"
            offs.append(len(out))
            out += sgm
        counts = [len(sgm) for sgm in segs]
        # arrays of offsets / counts
        def arr(vals):
            nonlocal out
            if len(out) % 2:
                out += b"\0"
            pos = len(out)
            out += b"".join(struct.pack(e + ("Q" if big else "I"), v) for v in vals)
            return pos
        po, pc = arr(offs), arr(counts)
        LONG = 16 if big else 4
        tags = [(256, 4, 1, W), (257, 4, 1, H), (258, 3, 1, bits),
                (259, 3, 1, 32773 if packbits else 1), (262, 3, 1, 1), (277, 3, 1, 1),
                (339, 3, 1, fmt)]
        if tile:
            tags += [(322, 4, 1, tile[1]), (323, 4, 1, tile[0]),
                     (324, LONG, len(offs), offs[0] if len(offs) == 1 else po),
                     (325, LONG, len(counts), counts[0] if len(counts) == 1 else pc)]
        else:
            tags += [(273, LONG, len(offs), offs[0] if len(offs) == 1 else po),
                     (278, 4, 1, rps or H),
                     (279, LONG, len(counts), counts[0] if len(counts) == 1 else pc)]
        tags.sort()
        if len(out) % 2:
            out += b"\0"
        ifd = len(out)
        struct.pack_into(e + ("Q" if big else "I"), out, prev_next_ptr, ifd)
        out += struct.pack(e + ("Q" if big else "H"), len(tags))
        for tag, typ, cnt, val in tags:
            if big:
                fmtv = {3: "H", 4: "I", 16: "Q"}[typ]
                v = struct.pack(e + fmtv, val).ljust(8, b"\0")
                out += struct.pack(e + "HHQ", tag, typ, cnt) + v
            else:
                fmtv = {3: "H", 4: "I"}[typ]
                v = struct.pack(e + fmtv, val).ljust(4, b"\0")
                out += struct.pack(e + "HHI", tag, typ, cnt) + v
        prev_next_ptr = len(out)
        out += struct.pack(e + ("Q" if big else "I"), 0)
    return bytes(out)


def ets_bytes(frames_by_channel, dims, code=4, extra=None, ndim=5):
    """The layout MinimalETSParser::new reads (decoderiter.rs:351-408):
    SIS header, ETS header at `addl`, chunk table at `used`; frames as u16.
    frames_by_channel: {channel: [2-D arrays]} (rows = dims[1], cols = dims[0])."""
    chunks = []
    data = bytearray()
    base = 64 + 64
    for ch, frs in sorted(frames_by_channel.items()):
        for t, fr in enumerate(frs):
            if fr is None:
                continue
            chunks.append((t, ch, base + len(data)))
            data += fr.astype("<u2").tobytes()
    if extra:
        chunks = extra(chunks)
    used = base + len(data)
    hdr = bytearray(64)
    hdr[0:4] = b"SIS\0"
    struct.pack_into("<I", hdr, 12, ndim)
    struct.pack_into("<Q", hdr, 16, 64)
    struct.pack_into("<Q", hdr, 32, used)
    struct.pack_into("<I", hdr, 40, len(chunks))
    ets = bytearray(64)
    ets[0:4] = b"ETS\0"
    struct.pack_into("<I", ets, 8, code)
    struct.pack_into("<I", ets, 28, dims[0])
    struct.pack_into("<I", ets, 32, dims[1])
    tab = bytearray()
    for t, ch, off in chunks:
        coords = [0] * ndim
        coords[2] = t
        coords[4] = ch
        tab += struct.pack("<I", 0) + struct.pack("<%dI" % ndim, *coords) + struct.pack("<QQ", off, 0)
    return bytes(hdr + ets + data + tab)
