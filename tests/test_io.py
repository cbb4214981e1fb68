"""Video-file input (path_to_iter, src/execute_gpu.rs:1186-1270): the
library's TIFF and ETS readers, checked on the CPU against Pillow's decoder
(an independent TIFF implementation) for every compression and sample type
it writes, against hand-built files for what Pillow does not write (tiles,
BigTIFF, big-endian pages, signed samples, the ETS container), and against
the reference's rules for errors, frame selection, the mean frame and
to_tiff's `f32 as u16` cast.  No GPU is needed: the readers are host code."""
import os
import struct
import zlib

import numpy as np
import pytest

import gpu_tracking_amd as gt
from gpu_tracking_amd import GtError

from filefmt import ets_bytes, tiff_bytes

PIL = pytest.importorskip("PIL.Image")


def _pil_pages(path):
    im = PIL.open(path)
    out = []
    i = 0
    while True:
        try:
            im.seek(i)
        except EOFError:
            break
        out.append(np.array(im))
        i += 1
    return np.stack(out)


def _status(excinfo):
    return excinfo.value.status


def _write(tmp_path, name, b):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(b)
    return p


# ---- Pillow cross-checks -----------------------------------------------------

@pytest.mark.parametrize("mode,dtype", [("L", np.uint8), ("I;16", np.uint16), ("F", np.float32)])
@pytest.mark.parametrize("comp", [None, "tiff_lzw", "packbits", "tiff_deflate",
                                  "tiff_adobe_deflate"])
def test_pillow_written_pages(tmp_path, mode, dtype, comp):
    rng = np.random.default_rng(zlib.crc32(f"{mode}/{comp}".encode()))
    if dtype == np.float32:
        arrs = [rng.normal(100, 30, (33, 47)).astype(np.float32) for _ in range(3)]
    else:
        arrs = [rng.integers(0, np.iinfo(dtype).max, (33, 47)).astype(dtype) for _ in range(3)]
    ims = [PIL.fromarray(a) for a in arrs]
    assert ims[0].mode == mode
    p = str(tmp_path / "v.tif")
    ims[0].save(p, save_all=True, append_images=ims[1:], compression=comp)
    ref = _pil_pages(p)
    assert np.array_equal(ref, np.stack(arrs))
    with gt.VideoFile(p) as f:
        assert f.shape == (3, 33, 47) and f.dtype == np.dtype(dtype)
        got = f.read()
    assert got.dtype == np.dtype(dtype) and np.array_equal(got, ref)


@pytest.mark.parametrize("kind", ["random", "ramp"])
@pytest.mark.parametrize("comp,pred", [("tiff_lzw", None), ("tiff_lzw", 2), ("tiff_deflate", 2),
                                       ("packbits", None)])
def test_pillow_large_frames_and_predictor(tmp_path, kind, comp, pred):
    # random data fills the LZW table (code widths 9..12 and clear codes)
    rng = np.random.default_rng(5)
    shape = (300, 1000)
    a = (rng.integers(0, 65535, shape) if kind == "random"
         else (np.add.outer(np.arange(shape[0]), np.arange(shape[1])) * 7) % 65536).astype(np.uint16)
    p = str(tmp_path / "b.tif")
    im = PIL.fromarray(a)
    if pred:
        im.save(p, compression=comp, tiffinfo={317: pred})
    else:
        im.save(p, compression=comp)
    ref = np.array(PIL.open(p))
    assert np.array_equal(ref, a)
    assert np.array_equal(gt.load(p)[0], a.astype(np.float32))


def test_big_endian_u16_from_pillow(tmp_path):
    a = np.random.default_rng(3).integers(0, 65535, (20, 31)).astype(np.uint16)
    im = PIL.frombytes("I;16B", (31, 20), a.astype(">u2").tobytes())
    p = str(tmp_path / "be.tif")
    im.save(p, compression="tiff_lzw")
    assert np.array_equal(gt.VideoFile(p).read()[0], a)


# ---- hand-built layouts ------------------------------------------------------

@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("be", [False, True])
@pytest.mark.parametrize("tile", [None, (16, 32)])
@pytest.mark.parametrize("packbits", [False, True])
def test_hand_built_layouts(tmp_path, big, be, tile, packbits):
    rng = np.random.default_rng(11)
    pages = [rng.integers(0, 65535, (45, 70)).astype(np.uint16) for _ in range(4)]
    pages[1][:, :30] = 7                     # runs for PackBits
    b = tiff_bytes(pages, big=big, be=be, tile=tile, rps=None if tile else 8, packbits=packbits)
    p = _write(tmp_path, "h.tif", b)
    with gt.VideoFile(p) as f:
        assert f.shape == (4, 45, 70)
        assert np.array_equal(f.read(), np.stack(pages))
        assert np.array_equal(f.read([3, 0, 2]), np.stack([pages[3], pages[0], pages[2]]))


@pytest.mark.parametrize("dtype", [np.int8, np.int16])
def test_signed_samples_read_as_f32(tmp_path, dtype):
    info = np.iinfo(dtype)
    a = np.random.default_rng(4).integers(info.min, info.max, (12, 9)).astype(dtype)
    p = _write(tmp_path, "s.tif", tiff_bytes([a, a[::-1].copy()], fmt=2))
    with gt.VideoFile(p) as f:
        assert f.dtype == np.float32
        got = f.read()
    assert np.array_equal(got, np.stack([a, a[::-1]]).astype(np.float32))


@pytest.mark.parametrize("dtype,fmt", [(np.uint32, 1), (np.int32, 2), (np.float64, 3)])
def test_wide_samples_are_a_cast_error(tmp_path, dtype, fmt):
    a = np.zeros((4, 4), dtype)
    p = _write(tmp_path, "w.tif", tiff_bytes([a], fmt=fmt))
    with pytest.raises(GtError) as e:
        gt.VideoFile(p)
    assert _status(e) == 14     # Error::CastError


def test_pages_of_different_sizes(tmp_path):
    b = bytearray(tiff_bytes([np.zeros((4, 4), np.uint16), np.zeros((4, 4), np.uint16)]))
    # the second page's width tag (256) says 5
    idx = b.rfind(struct.pack("<HHI", 256, 4, 1))
    struct.pack_into("<I", b, idx + 8, 5)
    p = _write(tmp_path, "d.tif", bytes(b))
    with pytest.raises(GtError) as e:
        gt.VideoFile(p)
    assert _status(e) == 4      # Error::DimensionMismatch


# ---- to_tiff and load --------------------------------------------------------

def test_to_tiff_casts_like_rust_and_round_trips(tmp_path):
    a = np.array([[[-5.0, 0.4, 0.9999, 1.5], [65534.7, 65535.0, 70000.0, np.nan],
                   [np.inf, -np.inf, 123.99, 7.0]]], np.float32)
    want = np.array([[[0, 0, 0, 1], [65534, 65535, 65535, 0], [65535, 0, 123, 7]]], np.uint16)
    p = str(tmp_path / "c.tif")
    gt.to_tiff(p, a)
    assert np.array_equal(gt.VideoFile(p).read(), want)
    assert np.array_equal(_pil_pages(p), want)
    assert gt.load(p).dtype == np.float32


def test_to_tiff_non_square_keeps_orientation(tmp_path):
    a = np.arange(2 * 5 * 9, dtype=np.float32).reshape(2, 5, 9)
    p = str(tmp_path / "n.tif")
    gt.to_tiff(p, a)
    ref = _pil_pages(p)
    assert ref.shape == (2, 5, 9) and np.array_equal(ref, a.astype(np.uint16))


def test_load_frames_selection_and_bounds(tmp_path):
    a = np.random.default_rng(8).integers(0, 1000, (6, 10, 12)).astype(np.float32)
    p = str(tmp_path / "l.tif")
    gt.to_tiff(p, a)
    assert np.array_equal(gt.load(p, frames=[4, 1, 1, 5]), a[[4, 1, 1, 5]])
    assert np.array_equal(gt.load(p, frames=range(3)), a[:3])
    with pytest.raises(GtError) as e:
        gt.load(p, frames=[2, 6])
    assert _status(e) == 7      # Error::FrameOutOfBounds


def test_mean_from_file_is_the_sequential_f32_mean(tmp_path):
    # large values: the f32 running sum rounds, so the order is visible
    a = np.random.default_rng(9).integers(60000, 65535, (300, 6, 7)).astype(np.float32)
    p = str(tmp_path / "m.tif")
    gt.to_tiff(p, a)
    acc = np.zeros((6, 7), np.float32)
    for t in range(a.shape[0]):
        acc = (acc + a[t]).astype(np.float32)
    want = (acc / np.float32(a.shape[0])).astype(np.float32)
    got = gt.mean_from_file(p)
    assert got.dtype == np.float32 and np.array_equal(got, want)
    # the float64 mean differs: the f32 running sum has rounded
    assert not np.array_equal(got, (a.astype(np.float64).mean(0)).astype(np.float32))


# ---- ETS / VSI -----------------------------------------------------------------

def test_ets_channels_frames_and_vsi(tmp_path):
    rng = np.random.default_rng(12)
    dims = (21, 13)   # frame shape (13, 21)
    c0 = [rng.integers(0, 65535, (13, 21)).astype(np.uint16) for _ in range(4)]
    c1 = [rng.integers(0, 65535, (13, 21)).astype(np.uint16) for _ in range(2)]
    b = ets_bytes({0: c0, 1: c1}, dims)
    p = _write(tmp_path, "v.ets", b)
    with gt.VideoFile(p) as f:
        assert f.shape == (4, 13, 21) and f.dtype == np.uint16
        assert f.channels() == [0, 1]
        assert np.array_equal(f.read(), np.stack(c0))
    assert np.array_equal(gt.VideoFile(p, channel=1).read(), np.stack(c1))
    d = gt.parse_ets(p)
    assert sorted(d) == [0, 1] and np.array_equal(d[1], np.stack(c1))
    assert np.array_equal(gt.parse_ets_with_keys(p, [3, 0], channel=0), np.stack([c0[3], c0[0]]))
    # x.vsi reads _x_/stack1/frame_t_0.ets beside it
    os.makedirs(tmp_path / "_x_" / "stack1")
    _write(tmp_path / "_x_" / "stack1", "frame_t_0.ets", b)
    vsi = _write(tmp_path, "x.vsi", b"")
    assert np.array_equal(gt.load(vsi, channel=1), np.stack(c1).astype(np.float32))
    with pytest.raises(GtError) as e:
        gt.VideoFile(p, channel=2)
    assert _status(e) == 17     # Error::ChannelNotFound


def test_ets_chunk_table_rules(tmp_path):
    # frames listed out of order, and a frame without a chunk: the table is
    # built with the reference's three cases; the hole reads as ReadError
    rng = np.random.default_rng(13)
    frs = [rng.integers(0, 65535, (5, 6)).astype(np.uint16) for _ in range(4)]
    frs[2] = None
    def reorder(chunks):
        return [chunks[2], chunks[0], chunks[1]]   # t = 3, 0, 1
    b = ets_bytes({0: frs}, (6, 5), extra=reorder)
    p = _write(tmp_path, "o.ets", b)
    with gt.VideoFile(p) as f:
        assert f.shape[0] == 4
        assert np.array_equal(f.read([0, 1, 3]), np.stack([frs[0], frs[1], frs[3]]))
        with pytest.raises(GtError) as e:
            f.read([2])
        assert _status(e) == 15     # Error::ReadError


def test_ets_other_data_types_are_rejected(tmp_path):
    b = ets_bytes({0: [np.zeros((3, 4), np.uint16)]}, (4, 3), code=2)
    with pytest.raises(GtError) as e:
        gt.VideoFile(_write(tmp_path, "u8.ets", b))
    assert _status(e) == 14


# ---- path errors ---------------------------------------------------------------

def test_path_errors(tmp_path):
    with pytest.raises(GtError) as e:
        gt.VideoFile(str(tmp_path / "noext"))
    assert _status(e) == 11     # NoExtensionError
    with pytest.raises(GtError) as e:
        gt.VideoFile(str(tmp_path / "a.png"))
    assert _status(e) == 12     # UnsupportedFileformat
    with pytest.raises(GtError) as e:
        gt.VideoFile(str(tmp_path / "missing.TIF"))
    assert _status(e) == 3      # FileNotFound
    with pytest.raises(GtError) as e:
        gt.VideoFile(_write(tmp_path, "bad.tif", b"XX*\0\0\0\0\0"))
    assert _status(e) == 15     # ReadError


# ---- malformed files end as ReadError, never as a crash --------------------

@pytest.mark.parametrize("be", [False, True])
@pytest.mark.parametrize("tw,th", [(2 ** 31, 2 ** 31), (2 ** 31, 16), (0, 16), (4096, 4096)])
def test_oversized_or_empty_tiles_are_read_errors(tmp_path, be, tw, th):
    # a 1x1 f32 image in one 16x16 tile, whose tile-size tags are then
    # rewritten (tw * th * 4 wraps to 0 in 64 bits for 2^31 x 2^31)
    e = ">" if be else "<"
    b = bytearray(tiff_bytes([np.ones((1, 1), np.float32)], be=be, tile=(16, 16), fmt=3))
    for tag, val in ((322, tw), (323, th)):
        idx = b.find(struct.pack(e + "HHI", tag, 4, 1))
        assert idx > 0
        struct.pack_into(e + "I", b, idx + 8, val)
    p = _write(tmp_path, "t.tif", bytes(b))
    with pytest.raises(ValueError) as err:
        gt.VideoFile(p)
    assert _status(err) == 15   # Error::ReadError


def test_ets_time_index_beyond_the_table_is_a_read_error(tmp_path):
    frs = [np.full((5, 6), 7, np.uint16)]

    def far(chunks):
        return [(2 ** 31, ch, off) for (t, ch, off) in chunks]

    p = _write(tmp_path, "t.ets", ets_bytes({0: frs}, (6, 5), extra=far))
    with pytest.raises(ValueError) as err:
        gt.VideoFile(p)
    assert _status(err) == 15


def test_ets_chunk_count_beyond_the_file_is_a_read_error(tmp_path):
    b = bytearray(ets_bytes({0: [np.zeros((5, 6), np.uint16)]}, (6, 5)))
    struct.pack_into("<I", b, 40, 0xFFFFFFFF)   # used-chunk count
    p = _write(tmp_path, "n.ets", bytes(b))
    with pytest.raises(ValueError) as err:
        gt.VideoFile(p)
    assert _status(err) == 15


def test_ets_frame_size_overflow_is_a_read_error(tmp_path):
    b = bytearray(ets_bytes({0: [np.zeros((5, 6), np.uint16)]}, (6, 5)))
    struct.pack_into("<II", b, 64 + 28, 0xFFFFFFFF, 0xFFFFFFFF)
    p = _write(tmp_path, "s.ets", bytes(b))
    with pytest.raises(ValueError) as err:
        gt.VideoFile(p)
    assert _status(err) == 15
