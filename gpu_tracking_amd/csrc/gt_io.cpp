// gt_io.cpp -- TIFF and ETS readers (see gt_io.h).
//
// TIFF: classic and BigTIFF, either byte order, strips or tiles, one sample
// per pixel, compression none / PackBits / LZW / Deflate, horizontal
// predictor; samples u8, u16, i8, i16, f32 (the types the reference decodes,
// decoderiter.rs:79-91; u32 / u64 / i32 / i64 / f64 are its CastError).
// Pages are indexed once at open; every page must have page 0's size and
// sample type.  Uncompressed little-endian strips are read with one pread
// straight into the destination.
//
// ETS: the reference's MinimalETSParser (decoderiter.rs:351-408): the SIS
// header, the ETS header's data type and frame size, and the used-chunk
// table keyed by channel (coordinate 4) and frame (coordinate 2), built with
// the same rules; frames are read as dims[0] * dims[1] little-endian u16
// (ETSIterator::next, :444-481).
#include "gt_io.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_set>

struct gt_file {
    int fd = -1;
    int kind = 0;            // 0 TIFF, 1 ETS
    bool be = false;         // TIFF byte order
    uint32_t H = 0, W = 0;
    int32_t dtype = GT_U16;  // pipeline dtype of a decoded frame
    int src = 0;             // sample type in the file (Sample below)
    struct Page {
        uint16_t compression = 1, predictor = 1;
        bool tiled = false;
        uint32_t rps = 0, tw = 0, th = 0;
        std::vector<uint64_t> offs, counts;
    };
    std::vector<Page> pages;
    std::map<uint32_t, std::vector<int64_t>> ets_all;  // channel -> offsets (-1: missing)
    std::vector<int64_t> ets;                          // the open channel's offsets
    uint64_t ets_bytes = 0;
    ~gt_file() {
        if (fd >= 0) close(fd);
    }
};

namespace {

enum Sample { S_U8, S_U16, S_I8, S_I16, S_F32 };

size_t sample_bytes(int s) { return s == S_U8 || s == S_I8 ? 1 : s == S_F32 ? 4 : 2; }

bool pread_all(int fd, uint64_t off, void *dst, size_t n) {
    char *p = static_cast<char *>(dst);
    while (n) {
        const ssize_t r = pread(fd, p, n, static_cast<off_t>(off));
        if (r <= 0) return false;
        p += r;
        off += static_cast<uint64_t>(r);
        n -= static_cast<size_t>(r);
    }
    return true;
}

uint16_t ld16(const uint8_t *p, bool be) {
    return be ? static_cast<uint16_t>(p[0] << 8 | p[1]) : static_cast<uint16_t>(p[1] << 8 | p[0]);
}
uint32_t ld32(const uint8_t *p, bool be) {
    return be ? (uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3])
              : (uint32_t(p[3]) << 24 | uint32_t(p[2]) << 16 | uint32_t(p[1]) << 8 | p[0]);
}
uint64_t ld64(const uint8_t *p, bool be) {
    return be ? (uint64_t(ld32(p, true)) << 32 | ld32(p + 4, true))
              : (uint64_t(ld32(p + 4, false)) << 32 | ld32(p, false));
}

// ---- TIFF ------------------------------------------------------------------

struct Entry {
    uint16_t tag = 0, type = 0;
    uint64_t count = 0;
    uint8_t raw[8] = {0};
};

size_t type_size(uint16_t t) {
    switch (t) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 0;
    }
}

// The integer values of an entry (BYTE / SHORT / LONG / LONG8 / IFD types).
bool entry_values(int fd, bool be, bool big, const Entry &e, std::vector<uint64_t> *v) {
    const size_t ts = type_size(e.type);
    // tag arrays hold at most one entry per strip or tile; 2^24 bounds any
    // real image and keeps a malformed count from a huge allocation
    if (!ts || e.count > (1ull << 24)) return false;
    const size_t n = static_cast<size_t>(e.count) * ts;
    std::vector<uint8_t> buf(n);
    const size_t inl = big ? 8 : 4;
    if (n <= inl) {
        std::memcpy(buf.data(), e.raw, n);
    } else {
        const uint64_t off = big ? ld64(e.raw, be) : ld32(e.raw, be);
        if (!pread_all(fd, off, buf.data(), n)) return false;
    }
    v->resize(static_cast<size_t>(e.count));
    for (size_t i = 0; i < v->size(); ++i) {
        const uint8_t *p = buf.data() + i * ts;
        switch (e.type) {
        case 1: case 7: (*v)[i] = p[0]; break;
        case 3: (*v)[i] = ld16(p, be); break;
        case 4: case 13: (*v)[i] = ld32(p, be); break;
        case 16: case 18: (*v)[i] = ld64(p, be); break;
        default: return false;
        }
    }
    return true;
}

gt_status open_tiff(gt_file *f, std::string *why) {
    uint8_t hdr[16];
    if (!pread_all(f->fd, 0, hdr, 8)) {
        *why = "could not read the TIFF header";
        return GT_ERR_READ;
    }
    if (hdr[0] == 'I' && hdr[1] == 'I') f->be = false;
    else if (hdr[0] == 'M' && hdr[1] == 'M') f->be = true;
    else {
        *why = "not a TIFF file (byte-order mark)";
        return GT_ERR_READ;
    }
    const bool be = f->be;
    const uint16_t magic = ld16(hdr + 2, be);
    bool big = false;
    uint64_t ifd = 0;
    if (magic == 42) {
        ifd = ld32(hdr + 4, be);
    } else if (magic == 43) {
        big = true;
        if (!pread_all(f->fd, 8, hdr + 8, 8) || ld16(hdr + 4, be) != 8) {
            *why = "malformed BigTIFF header";
            return GT_ERR_READ;
        }
        ifd = ld64(hdr + 8, be);
    } else {
        *why = "not a TIFF file (magic number)";
        return GT_ERR_READ;
    }
    int src0 = -1;
    std::unordered_set<uint64_t> seen;
    while (ifd != 0) {
        if (!seen.insert(ifd).second) break;   // a cycle in the IFD chain
        uint8_t cb[8];
        const size_t csz = big ? 8 : 2, esz = big ? 20 : 12;
        if (!pread_all(f->fd, ifd, cb, csz)) {
            *why = "could not read an image file directory";
            return GT_ERR_READ;
        }
        const uint64_t n = big ? ld64(cb, be) : ld16(cb, be);
        if (n == 0 || n > 65536) {
            *why = "malformed image file directory";
            return GT_ERR_READ;
        }
        std::vector<uint8_t> eb(static_cast<size_t>(n) * esz + (big ? 8 : 4));
        if (!pread_all(f->fd, ifd + csz, eb.data(), eb.size())) {
            *why = "could not read an image file directory";
            return GT_ERR_READ;
        }
        uint64_t W = 0, H = 0, bits = 8, fmt = 1, spp = 1, planar = 1, fill = 1;
        gt_file::Page pg;
        pg.rps = 0xFFFFFFFFu;
        std::vector<uint64_t> v;
        for (uint64_t i = 0; i < n; ++i) {
            const uint8_t *p = eb.data() + i * esz;
            Entry e;
            e.tag = ld16(p, be);
            e.type = ld16(p + 2, be);
            e.count = big ? ld64(p + 4, be) : ld32(p + 4, be);
            std::memcpy(e.raw, p + (big ? 12 : 8), big ? 8 : 4);
            const bool wanted = (e.tag >= 256 && e.tag <= 259) || e.tag == 266 || e.tag == 273 ||
                                e.tag == 277 || e.tag == 278 || e.tag == 279 || e.tag == 284 ||
                                e.tag == 317 || (e.tag >= 322 && e.tag <= 325) || e.tag == 339;
            if (!wanted) continue;
            if (!entry_values(f->fd, be, big, e, &v) || v.empty()) {
                *why = "malformed TIFF tag " + std::to_string(e.tag);
                return GT_ERR_READ;
            }
            switch (e.tag) {
            case 256: W = v[0]; break;
            case 257: H = v[0]; break;
            case 258: bits = v[0]; break;
            case 259: pg.compression = static_cast<uint16_t>(v[0]); break;
            case 266: fill = v[0]; break;
            case 273: pg.offs = v; break;
            case 277: spp = v[0]; break;
            case 278: pg.rps = static_cast<uint32_t>(std::min<uint64_t>(v[0], 0xFFFFFFFFu)); break;
            case 279: pg.counts = v; break;
            case 284: planar = v[0]; break;
            case 317: pg.predictor = static_cast<uint16_t>(v[0]); break;
            case 322: pg.tw = static_cast<uint32_t>(v[0]); pg.tiled = true; break;
            case 323: pg.th = static_cast<uint32_t>(v[0]); pg.tiled = true; break;
            case 324: pg.offs = v; pg.tiled = true; break;
            case 325: pg.counts = v; pg.tiled = true; break;
            case 339: fmt = v[0]; break;
            }
        }
        const uint8_t *np = eb.data() + static_cast<size_t>(n) * esz;
        const uint64_t next = big ? ld64(np, be) : ld32(np, be);
        if (spp != 1 || planar != 1 || fill != 1) {
            *why = "only single-sample (grey) TIFF images are supported";
            return GT_ERR_READ;
        }
        int src;
        if (fmt == 1 && bits == 8) src = S_U8;
        else if (fmt == 1 && bits == 16) src = S_U16;
        else if (fmt == 2 && bits == 8) src = S_I8;
        else if (fmt == 2 && bits == 16) src = S_I16;
        else if (fmt == 3 && bits == 32) src = S_F32;
        else if ((fmt == 1 || fmt == 2) && (bits == 32 || bits == 64)) {
            *why = "Could not cast the image data to f32 (integer samples of 32 or 64 bits)";
            return GT_ERR_CAST;
        } else if (fmt == 3 && bits == 64) {
            *why = "Could not cast the image data to f32 (f64 samples)";
            return GT_ERR_CAST;
        } else {
            *why = "unsupported TIFF sample format " + std::to_string(fmt) + " with " +
                   std::to_string(bits) + " bits";
            return GT_ERR_READ;
        }
        // a frame is decoded into one buffer: its byte size must fit far
        // below size_t (2^36 samples bounds any real frame)
        if (W == 0 || H == 0 || W > 0xFFFFFFFFull || H > 0xFFFFFFFFull ||
            W * H > (1ull << 36)) {
            *why = "TIFF image without a valid size";
            return GT_ERR_READ;
        }
        const uint16_t c = pg.compression;
        if (c != 1 && c != 5 && c != 8 && c != 32946 && c != 32773) {
            *why = "unsupported TIFF compression " + std::to_string(c);
            return GT_ERR_READ;
        }
        if (pg.predictor != 1 && !(pg.predictor == 2 && src != S_F32)) {
            *why = "unsupported TIFF predictor " + std::to_string(pg.predictor);
            return GT_ERR_READ;
        }
        uint64_t nseg;
        if (pg.tiled) {
            if (pg.tw == 0 || pg.th == 0) {
                *why = "tiled TIFF without a tile size";
                return GT_ERR_READ;
            }
            // one tile is decoded into a buffer of tw * th samples: tiles
            // larger than the image rounded up to a multiple of 256 (or
            // above 2^28 samples) are malformed, and this bounds the
            // product far from overflow
            const uint64_t wmax = (W + 255) / 256 * 256, hmax = (H + 255) / 256 * 256;
            if (pg.tw > wmax || pg.th > hmax ||
                static_cast<uint64_t>(pg.tw) * pg.th > (1ull << 28)) {
                *why = "TIFF tile of " + std::to_string(pg.tw) + " x " + std::to_string(pg.th) +
                       " for a " + std::to_string(W) + " x " + std::to_string(H) + " image";
                return GT_ERR_READ;
            }
            nseg = ((W + pg.tw - 1) / pg.tw) * ((H + pg.th - 1) / pg.th);
        } else {
            if (pg.rps == 0) pg.rps = 1;
            nseg = (H + pg.rps - 1) / pg.rps;
        }
        if (pg.offs.size() < nseg || pg.counts.size() < nseg) {
            *why = "TIFF strip or tile table shorter than the image";
            return GT_ERR_READ;
        }
        if (f->pages.empty()) {
            f->W = static_cast<uint32_t>(W);
            f->H = static_cast<uint32_t>(H);
            src0 = src;
        } else if (W != f->W || H != f->H || src != src0) {
            *why = "page " + std::to_string(f->pages.size()) +
                   " differs from the first page in size or sample type";
            return GT_ERR_DIMENSION_MISMATCH;
        }
        f->pages.push_back(std::move(pg));
        ifd = next;
    }
    if (f->pages.empty()) {
        *why = "the TIFF file holds no image";
        return GT_ERR_EMPTY_ITERATOR;
    }
    f->src = src0;
    f->dtype = src0 == S_U8 ? GT_U8 : src0 == S_U16 ? GT_U16 : GT_F32;
    return GT_OK;
}

bool unpack_bits(const uint8_t *in, size_t n, uint8_t *out, size_t want) {
    size_t i = 0, o = 0;
    while (i < n && o < want) {
        const int8_t h = static_cast<int8_t>(in[i++]);
        if (h >= 0) {
            const size_t k = static_cast<size_t>(h) + 1;
            if (i + k > n || o + k > want) return false;
            std::memcpy(out + o, in + i, k);
            i += k;
            o += k;
        } else if (h != -128) {
            const size_t k = static_cast<size_t>(1 - h);
            if (i >= n || o + k > want) return false;
            std::memset(out + o, in[i++], k);
            o += k;
        }
    }
    return o == want;
}

// TIFF LZW: MSB-first codes of 9..12 bits, 256 = clear, 257 = end; the
// code width grows one code early (when the next free code reaches 2^w - 1).
bool unlzw(const uint8_t *in, size_t n, uint8_t *out, size_t want) {
    std::vector<uint16_t> prefix(4096), len(4096);
    std::vector<uint8_t> suffix(4096), first(4096);
    for (int i = 0; i < 256; ++i) {
        suffix[i] = static_cast<uint8_t>(i);
        first[i] = static_cast<uint8_t>(i);
        len[i] = 1;
    }
    size_t bitpos = 0, o = 0;
    int width = 9, next = 258, prev = -1;
    const size_t nbits = n * 8;
    auto emit = [&](int code) -> bool {
        const size_t l = len[code];
        if (o + l > want) return false;
        size_t pos = o + l;
        int c = code;
        for (size_t k = 0; k < l; ++k) {
            out[--pos] = suffix[c];
            c = prefix[c];
        }
        o += l;
        return true;
    };
    while (o < want) {
        if (bitpos + width > nbits) return false;
        int code = 0;
        for (int b = 0; b < width; ++b, ++bitpos)
            code = code << 1 | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255 || !emit(code)) return false;
            prev = code;
            continue;
        }
        int fc;
        if (code < next) fc = first[code];
        else if (code == next) fc = first[prev];   // the string being defined
        else return false;
        if (next < 4096) {
            prefix[next] = static_cast<uint16_t>(prev);
            suffix[next] = static_cast<uint8_t>(fc);
            first[next] = first[prev];
            len[next] = static_cast<uint16_t>(len[prev] + 1);
            ++next;
        }
        if (!emit(code)) return false;
        prev = code;
        if (next >= (1 << width) - 1 && width < 12) ++width;
    }
    return o == want;
}

bool inflate_all(const uint8_t *in, size_t n, uint8_t *out, size_t want) {
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef *>(in);
    zs.avail_in = static_cast<uInt>(n);
    zs.next_out = out;
    zs.avail_out = static_cast<uInt>(want);
    const int r = inflate(&zs, Z_FINISH);
    const size_t got = want - zs.avail_out;
    inflateEnd(&zs);
    return (r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR) && got == want;
}

// Decodes one strip or tile of `want` bytes (file byte order) into out.
bool decode_segment(const gt_file *f, const gt_file::Page &pg, size_t seg, uint8_t *out,
                    size_t want, std::vector<uint8_t> &tmp) {
    const uint64_t off = pg.offs[seg], cnt = pg.counts[seg];
    if (pg.compression == 1) {
        if (cnt < want) return false;
        return pread_all(f->fd, off, out, want);
    }
    if (cnt > (static_cast<uint64_t>(want) + 1024) * 4 + (1u << 20)) return false;   // implausible
    tmp.resize(static_cast<size_t>(cnt));
    if (!pread_all(f->fd, off, tmp.data(), tmp.size())) return false;
    switch (pg.compression) {
    case 32773: return unpack_bits(tmp.data(), tmp.size(), out, want);
    case 5: return unlzw(tmp.data(), tmp.size(), out, want);
    case 8: case 32946: return inflate_all(tmp.data(), tmp.size(), out, want);
    default: return false;
    }
}

// Byte order and the horizontal predictor over rows x cols samples.
void fix_samples(uint8_t *p, size_t rows, size_t cols, size_t stride, int src, bool swap,
                 uint16_t predictor) {
    const size_t bps = sample_bytes(src);
    for (size_t r = 0; r < rows; ++r) {
        uint8_t *row = p + r * stride * bps;
        if (swap && bps > 1) {
            for (size_t i = 0; i < cols; ++i) std::reverse(row + i * bps, row + (i + 1) * bps);
        }
        if (predictor == 2) {
            if (bps == 1) {
                for (size_t i = 1; i < cols; ++i) row[i] = static_cast<uint8_t>(row[i] + row[i - 1]);
            } else {
                uint16_t *q = reinterpret_cast<uint16_t *>(row);
                for (size_t i = 1; i < cols; ++i) q[i] = static_cast<uint16_t>(q[i] + q[i - 1]);
            }
        }
    }
}

gt_status read_tiff(const gt_file *f, uint64_t idx, void *dst, std::string *why) {
    const gt_file::Page &pg = f->pages[static_cast<size_t>(idx)];
    const size_t W = f->W, H = f->H, bps = sample_bytes(f->src);
    const bool widen = f->src == S_I8 || f->src == S_I16;
    std::vector<uint8_t> img;
    uint8_t *base = static_cast<uint8_t *>(dst);
    if (widen) {
        img.resize(W * H * bps);
        base = img.data();
    }
    std::vector<uint8_t> tmp, tile;
    const bool swap = f->be;
    if (!pg.tiled) {
        const size_t nseg = (H + pg.rps - 1) / pg.rps;
        for (size_t s = 0; s < nseg; ++s) {
            const size_t r0 = s * pg.rps, rows = std::min<size_t>(pg.rps, H - r0);
            uint8_t *o = base + r0 * W * bps;
            if (!decode_segment(f, pg, s, o, rows * W * bps, tmp)) {
                *why = "could not decode strip " + std::to_string(s) + " of frame " + std::to_string(idx);
                return GT_ERR_READ;
            }
            fix_samples(o, rows, W, W, f->src, swap, pg.predictor);
        }
    } else {
        const size_t tw = pg.tw, th = pg.th, ta = (W + tw - 1) / tw, td = (H + th - 1) / th;
        tile.resize(tw * th * bps);
        for (size_t ty = 0; ty < td; ++ty) {
            for (size_t tx = 0; tx < ta; ++tx) {
                const size_t s = ty * ta + tx;
                if (!decode_segment(f, pg, s, tile.data(), tile.size(), tmp)) {
                    *why = "could not decode tile " + std::to_string(s) + " of frame " + std::to_string(idx);
                    return GT_ERR_READ;
                }
                fix_samples(tile.data(), th, tw, tw, f->src, swap, pg.predictor);
                const size_t rows = std::min(th, H - ty * th), cols = std::min(tw, W - tx * tw);
                for (size_t r = 0; r < rows; ++r)
                    std::memcpy(base + ((ty * th + r) * W + tx * tw) * bps, tile.data() + r * tw * bps,
                                cols * bps);
            }
        }
    }
    if (widen) {
        float *o = static_cast<float *>(dst);
        if (f->src == S_I8) {
            for (size_t i = 0; i < W * H; ++i) o[i] = static_cast<float>(static_cast<int8_t>(img[i]));
        } else {
            const int16_t *q = reinterpret_cast<const int16_t *>(img.data());
            for (size_t i = 0; i < W * H; ++i) o[i] = static_cast<float>(q[i]);
        }
    }
    return GT_OK;
}

// ---- ETS -------------------------------------------------------------------

gt_status open_ets(gt_file *f, int32_t has_channel, uint32_t channel, std::string *why) {
    auto rd = [&](uint64_t off, void *p, size_t n) { return pread_all(f->fd, off, p, n); };
    uint8_t b[64];
    if (!rd(0, b, 36) || std::memcmp(b, "SIS\0", 4) != 0) {
        *why = "not an ETS file (SIS header)";
        return GT_ERR_READ;
    }
    // SIS header: magic, 8 bytes, ndims u32 @12, additional header u64 @16,
    // 8 bytes, used-chunk table u64 @32... (decoderiter.rs:352-364)
    const uint32_t ndim = ld32(b + 12, false);
    const uint64_t addl = ld64(b + 16, false);
    if (!rd(32, b, 12)) {
        *why = "truncated ETS header";
        return GT_ERR_READ;
    }
    const uint64_t used = ld64(b, false);
    const uint32_t nused = ld32(b + 8, false);
    if (ndim < 5 || ndim > 16) {
        *why = "ETS chunks with fewer than 5 coordinates";
        return GT_ERR_READ;
    }
    if (!rd(addl, b, 36) || std::memcmp(b, "ETS\0", 4) != 0) {
        *why = "not an ETS file (ETS header)";
        return GT_ERR_READ;
    }
    const uint32_t code = ld32(b + 8, false);
    if (code != 4) {
        *why = code == 1 || code == 2 || code == 3 || code == 5 || code == 6 || code == 9
                   ? "only u16 ETS data is supported (data type code " + std::to_string(code) + ")"
                   : "Unknown ETS data type code: " + std::to_string(code);
        return code == 1 || code == 2 || code == 3 || code == 5 || code == 6 || code == 9 ? GT_ERR_CAST
                                                                                           : GT_ERR_READ;
    }
    const uint32_t d0 = ld32(b + 28, false), d1 = ld32(b + 32, false);
    if (d0 == 0 || d1 == 0 || static_cast<uint64_t>(d0) * d1 > (1ull << 36)) {
        *why = "ETS frame of size " + std::to_string(d0) + " x " + std::to_string(d1);
        return GT_ERR_READ;
    }
    const size_t esz = 4 + 4 * static_cast<size_t>(ndim) + 8 + 8;
    // the chunk table lies in the file: its length bounds the entry count
    struct stat stt;
    if (fstat(f->fd, &stt) != 0 || static_cast<uint64_t>(nused) * esz > static_cast<uint64_t>(stt.st_size)) {
        *why = "ETS chunk table of " + std::to_string(nused) + " entries exceeds the file";
        return GT_ERR_READ;
    }
    std::vector<uint8_t> tab(esz * nused);
    if (nused && !rd(used, tab.data(), tab.size())) {
        *why = "truncated ETS chunk table";
        return GT_ERR_READ;
    }
    for (uint32_t i = 0; i < nused; ++i) {
        const uint8_t *p = tab.data() + esz * i + 4;
        const uint32_t t = ld32(p + 4 * 2, false), ch = ld32(p + 4 * 4, false);
        const int64_t off = static_cast<int64_t>(ld64(p + 4 * ndim, false));
        // frames missing from the table read as errors later (the reference
        // keeps such holes); the table of a channel stays within 2^20
        // entries beyond the file's chunk count
        if (static_cast<uint64_t>(t) >= static_cast<uint64_t>(nused) + (1u << 20)) {
            *why = "ETS chunk " + std::to_string(i) + " has time index " + std::to_string(t) +
                   " beyond the " + std::to_string(nused) + " chunks of the file";
            return GT_ERR_READ;
        }
        std::vector<int64_t> &e = f->ets_all[ch];
        // the reference's three cases (decoderiter.rs:388-396)
        if (static_cast<size_t>(t) + 1 == e.size()) {
            e.push_back(off);
        } else if (static_cast<size_t>(t) + 1 > e.size()) {
            e.resize(static_cast<size_t>(t) + 1, -1);
            e[t] = off;
        } else {
            e[t] = off;
        }
    }
    const uint32_t want = has_channel ? channel : 0;
    auto it = f->ets_all.find(want);
    if (it == f->ets_all.end()) {
        *why = "channel " + std::to_string(want) + " not found in the ETS file";
        return GT_ERR_CHANNEL_NOT_FOUND;
    }
    f->ets = it->second;
    // frames are (dims[1], dims[0]) (execute_gpu.rs:1228)
    f->H = d1;
    f->W = d0;
    f->src = S_U16;
    f->dtype = GT_U16;
    f->ets_bytes = static_cast<uint64_t>(d0) * d1 * 2;
    if (f->ets.empty()) {
        *why = "the ETS channel holds no frame";
        return GT_ERR_EMPTY_ITERATOR;
    }
    return GT_OK;
}

std::string lower_ext(const std::string &path, std::string *stem, std::string *dir) {
    const size_t slash = path.find_last_of('/');
    const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    *dir = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
    const size_t dot = name.find_last_of('.');
    if (dot == std::string::npos || dot == 0) {
        *stem = name;
        return std::string();
    }
    *stem = name.substr(0, dot);
    std::string e = name.substr(dot + 1);
    for (char &ch : e) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    return e;
}

}  // namespace

namespace {

gt_status io_open(const char *path, int32_t has_channel, uint32_t channel, gt_file **out,
                  std::string *why) {
    *out = nullptr;
    std::string stem, dir;
    const std::string p(path);
    const std::string ext = lower_ext(p, &stem, &dir);
    if (ext.empty()) {
        *why = "No file extension found for " + p;
        return GT_ERR_NO_EXTENSION;
    }
    std::string open_path = p;
    int kind;
    if (ext == "tif" || ext == "tiff") {
        kind = 0;
    } else if (ext == "ets") {
        kind = 1;
    } else if (ext == "vsi") {
        // the frames live in _<stem>_/stack1/frame_t_0.ets beside the .vsi
        kind = 1;
        open_path = dir + "/_" + stem + "_/stack1/frame_t_0.ets";
    } else {
        *why = "Unsupported file format: " + ext;
        return GT_ERR_UNSUPPORTED_FILEFORMAT;
    }
    gt_file *f = new gt_file();
    f->kind = kind;
    f->fd = open(open_path.c_str(), O_RDONLY | O_CLOEXEC);
    if (f->fd < 0) {
        *why = "File not found: " + p + (open_path != p ? " (looked for " + open_path + ")" : "");
        delete f;
        return GT_ERR_FILE_NOT_FOUND;
    }
    const gt_status st = kind == 0 ? open_tiff(f, why) : open_ets(f, has_channel, channel, why);
    if (st != GT_OK) {
        delete f;
        return st;
    }
    *out = f;
    return GT_OK;
}

}  // namespace

// Sizes read from a file bound every allocation above, but a failed
// allocation must still end as an error, not cross the C boundary.
gt_status gt_io_open(const char *path, int32_t has_channel, uint32_t channel, gt_file **out,
                     std::string *why) {
    try {
        return io_open(path, has_channel, channel, out, why);
    } catch (const std::exception &e) {
        *out = nullptr;
        *why = std::string("could not open the file: ") + e.what();
        return GT_ERR_READ;
    }
}

void gt_io_close(gt_file *f) { delete f; }

uint64_t gt_io_nframes(const gt_file *f) { return f->kind == 0 ? f->pages.size() : f->ets.size(); }
int32_t gt_io_dtype(const gt_file *f) { return f->dtype; }
void gt_io_dims(const gt_file *f, uint32_t *nrows, uint32_t *ncols) {
    *nrows = f->H;
    *ncols = f->W;
}

gt_status gt_io_read(const gt_file *f, uint64_t idx, void *dst, std::string *why) {
    const uint64_t n = gt_io_nframes(f);
    if (idx >= n) {
        *why = "Requested a frame outside the length\n of the video. Video length is " +
               std::to_string(n) + ", requested frame " + std::to_string(idx);
        return GT_ERR_FRAME_OUT_OF_BOUNDS;
    }
    if (f->kind == 0) {
        try {
            return read_tiff(f, idx, dst, why);
        } catch (const std::exception &e) {
            *why = "could not decode frame " + std::to_string(idx) + ": " + e.what();
            return GT_ERR_READ;
        }
    }
    const int64_t off = f->ets[static_cast<size_t>(idx)];
    if (off < 0) {
        *why = "frame " + std::to_string(idx) + " has no chunk in the ETS file";
        return GT_ERR_READ;
    }
    if (!pread_all(f->fd, static_cast<uint64_t>(off), dst, static_cast<size_t>(f->ets_bytes))) {
        *why = "could not read frame " + std::to_string(idx) + " of the ETS file";
        return GT_ERR_READ;
    }
    return GT_OK;
}

std::vector<uint32_t> gt_io_channels(const gt_file *f) {
    std::vector<uint32_t> v;
    for (const auto &kv : f->ets_all) v.push_back(kv.first);
    return v;
}

namespace {

// Rust's saturating `f32 as u16`: NaN -> 0, truncation toward zero.
uint16_t as_u16(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 65535.0f) return 65535;
    return static_cast<uint16_t>(x);
}

void put16(std::vector<uint8_t> &b, uint16_t v) {
    b.push_back(static_cast<uint8_t>(v));
    b.push_back(static_cast<uint8_t>(v >> 8));
}
void put32(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(static_cast<uint8_t>(v >> (8 * i)));
}
void put64(std::vector<uint8_t> &b, uint64_t v) {
    for (int i = 0; i < 8; ++i) b.push_back(static_cast<uint8_t>(v >> (8 * i)));
}

}  // namespace

gt_status gt_io_write_tiff_u16(const char *path, const float *frames, uint64_t nframes,
                               uint32_t nrows, uint32_t ncols, std::string *why) {
    const uint64_t plane = 2ull * nrows * ncols;
    const int ntag = 10;
    const bool big = 16 + nframes * (plane + 8 + 20 * ntag + 8) >= 0xFFFFFFF0ull;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        *why = std::string("could not create ") + path;
        return GT_ERR_TIFF_WRITE;
    }
    std::vector<uint8_t> hdr;
    hdr.push_back('I');
    hdr.push_back('I');
    const uint64_t first_ifd = (big ? 16 : 8) + plane;
    if (big) {
        put16(hdr, 43);
        put16(hdr, 8);
        put16(hdr, 0);
        put64(hdr, nframes ? first_ifd : 0);
    } else {
        put16(hdr, 42);
        put32(hdr, nframes ? static_cast<uint32_t>(first_ifd) : 0);
    }
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size();
    uint64_t pos = hdr.size();
    std::vector<uint8_t> px(static_cast<size_t>(plane));
    const size_t ifd_bytes = big ? 8 + 20 * ntag + 8 : 2 + 12 * ntag + 4;
    for (uint64_t t = 0; t < nframes && ok; ++t) {
        const float *fr = frames + t * static_cast<uint64_t>(nrows) * ncols;
        for (size_t i = 0; i < static_cast<size_t>(nrows) * ncols; ++i) {
            const uint16_t v = as_u16(fr[i]);
            px[2 * i] = static_cast<uint8_t>(v);
            px[2 * i + 1] = static_cast<uint8_t>(v >> 8);
        }
        const uint64_t data_off = pos;
        ok = std::fwrite(px.data(), 1, px.size(), fp) == px.size();
        pos += plane;
        const uint64_t next = t + 1 < nframes ? pos + ifd_bytes + plane : 0;
        std::vector<uint8_t> ifd;
        auto entry = [&](uint16_t tag, uint16_t type, uint64_t value) {
            put16(ifd, tag);
            put16(ifd, type);
            if (big) {
                put64(ifd, 1);
                put64(ifd, value);
            } else {
                put32(ifd, 1);
                put32(ifd, static_cast<uint32_t>(value));
            }
        };
        if (big) put64(ifd, ntag); else put16(ifd, ntag);
        const uint16_t LONG = big ? 16 : 4;
        entry(256, 4, ncols);              // ImageWidth
        entry(257, 4, nrows);              // ImageLength
        entry(258, 3, 16);                 // BitsPerSample
        entry(259, 3, 1);                  // Compression: none
        entry(262, 3, 1);                  // Photometric: black is zero
        entry(273, LONG, data_off);        // StripOffsets
        entry(277, 3, 1);                  // SamplesPerPixel
        entry(278, 4, nrows);              // RowsPerStrip
        entry(279, LONG, plane);           // StripByteCounts
        entry(339, 3, 1);                  // SampleFormat: unsigned
        if (big) put64(ifd, next); else put32(ifd, static_cast<uint32_t>(next));
        ok = ok && std::fwrite(ifd.data(), 1, ifd.size(), fp) == ifd.size();
        pos += ifd.size();
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) {
        *why = std::string("could not write ") + path;
        return GT_ERR_TIFF_WRITE;
    }
    return GT_OK;
}
