"""File formats of libetn.so (csrc/io.cpp) on the CPU: TIFF images and
image-stack directories, MRC projection stacks and volumes, tilt files.

What pins them:
* TIFF: LZW files written by the reference's own FreeImage
  (tests/golden/files/projection-1.tif, projection60.tif) decode to the pixels
  Pillow decodes for the same files (synthetic_3_parallel.npz); Pillow also
  writes the other compressions and sample types the reader accepts.
* MRC volumes: tests/golden/files/voxelized_reference.mrc was written by the
  reference's MRC writer in its default x-z-y order; reading it and writing it
  again reproduces the file byte for byte, header included.
* MRC stacks: headers built here with numpy; the log transform is checked at
  its closed-form end points (raw minimum -> log(max - min + 1), raw maximum -> 0).
"""
import os
import struct
import zlib

import numpy as np
import pytest

from ettention_amd import io as EI
from ettention_amd._lib import InvalidArgument, Unsupported
from tests._util import GOLDEN, load_golden

FILES = os.path.join(GOLDEN, "files")
REF_DATA = "/root/reference/ettention/test/data"
needs_reference = pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference test data not present")


@pytest.fixture(scope="module")
def synthetic():
    return load_golden("synthetic_3_parallel")


# ---- TIFF -------------------------------------------------------------------

@pytest.mark.parametrize("angle", [-1, 60])
def test_lzw_tiffs_of_the_reference_decode_like_pillow(synthetic, angle):
    img = EI.read_tiff(os.path.join(FILES, "projection%d.tif" % angle))
    k = int(np.flatnonzero(synthetic["angles"] == angle)[0])
    assert img.dtype == np.float32 and img.shape == (256, 256)
    np.testing.assert_array_equal(img, synthetic["projections"][k])


@needs_reference
def test_reference_image_stack_directory(synthetic):
    proj, tilts, base = EI.open_projections(os.path.join(REF_DATA, "synthetic_3_parallel"))
    np.testing.assert_array_equal(proj, synthetic["projections"])
    np.testing.assert_array_equal(tilts, synthetic["angles"])
    # metadata.xml: detectorPosition -64 -64 128, sourcePosition -64 -64 -128, pitches 0.5
    assert list(base.source) == [-64, -64, -128] and list(base.detector) == [-64, -64, 128]
    assert list(base.horizontal_pitch) == [0.5, 0, 0] and list(base.vertical_pitch) == [0, 0.5, 0]
    assert list(base.resolution) == [256, 256]


PIL = pytest.importorskip("PIL.Image")


@pytest.mark.parametrize("compression", [None, "tiff_lzw", "tiff_adobe_deflate", "packbits"])
def test_float_tiff_compressions_against_pillow(tmp_path, compression):
    rng = np.random.default_rng(3)
    a = rng.normal(0, 100, (37, 53)).astype(np.float32)
    p = str(tmp_path / "a.tif")
    kw = {"compression": compression} if compression else {}
    PIL.fromarray(a, mode="F").save(p, **kw)
    np.testing.assert_array_equal(EI.read_tiff(p), np.array(PIL.open(p), np.float32))


@pytest.mark.parametrize("mode,scale", [("L", 255.0), ("I;16", 65535.0)])
def test_integer_tiffs_scale_to_unit_range(tmp_path, mode, scale):
    rng = np.random.default_rng(4)
    dt = np.uint8 if mode == "L" else np.uint16
    a = rng.integers(0, int(scale) + 1, (20, 31)).astype(dt)
    p = str(tmp_path / "a.tif")
    PIL.fromarray(a).save(p, compression="tiff_lzw")
    got = EI.read_tiff(p)
    np.testing.assert_array_equal(got, a.astype(np.float32) / np.float32(scale))
    assert got.max() <= 1.0


def test_multipage_tiff(tmp_path):
    pages = [np.full((8, 9), i + 0.5, np.float32) for i in range(3)]
    p = str(tmp_path / "stack.tif")
    PIL.fromarray(pages[0], mode="F").save(p, save_all=True,
                                           append_images=[PIL.fromarray(x, mode="F") for x in pages[1:]])
    assert EI.tiff_page_count(p) == 3
    for i in range(3):
        np.testing.assert_array_equal(EI.read_tiff(p, i), pages[i])
    with pytest.raises(InvalidArgument):
        EI.read_tiff(p, 3)


def _tiff_bytes(width, height, strips, bits, fmt, predictor, big_endian, compression=8):
    """A TIFF built from the TIFF 6.0 tag layout: one IFD, Deflate strips (the
    predictor applies to LZW and Deflate data only)."""
    e = ">" if big_endian else "<"
    rows_per_strip = -(-height // len(strips))
    strips = [zlib.compress(s) for s in strips] if compression == 8 else strips
    entries = [(256, 4, width), (257, 4, height), (258, 3, bits), (259, 3, compression), (262, 3, 1), (277, 3, 1),
               (278, 4, rows_per_strip), (317, 3, predictor), (339, 3, fmt)]
    n = len(entries) + 2
    data_off = 8 + 2 + 12 * n + 4 + 8 * len(strips)
    offs, o = [], data_off
    for s in strips:
        offs.append(o)
        o += len(s)
    arrays_off = 8 + 2 + 12 * n + 4
    entries += [(273, 4, ("arr", arrays_off, offs)), (279, 4, ("arr", arrays_off + 4 * len(strips),
                                                                   [len(s) for s in strips]))]
    entries.sort(key=lambda t: t[0])
    out = (b"MM" if big_endian else b"II") + struct.pack(e + "HI", 42, 8) + struct.pack(e + "H", len(entries))
    for tag, typ, val in entries:
        if isinstance(val, tuple):   # one LONG fits the entry itself, more are stored at an offset
            out += struct.pack(e + "HHII", tag, typ, len(val[2]), val[2][0] if len(val[2]) == 1 else val[1])
        elif typ == 3:
            out += struct.pack(e + "HHIHH", tag, typ, 1, val, 0)
        else:
            out += struct.pack(e + "HHII", tag, typ, 1, val)
    out += struct.pack(e + "I", 0)
    out += struct.pack(e + "%dI" % len(strips), *offs) + struct.pack(e + "%dI" % len(strips), *[len(s) for s in strips])
    assert len(out) == data_off
    return out + b"".join(strips)


@pytest.mark.parametrize("big_endian", [False, True])
def test_horizontal_predictor_uint16_in_strips(tmp_path, big_endian):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 65536, (7, 11)).astype(np.uint16)
    d = a.astype(np.int64)
    d[:, 1:] = (a[:, 1:].astype(np.int64) - a[:, :-1].astype(np.int64)) % 65536   # differenced along rows
    raw = d.astype(">u2" if big_endian else "<u2").tobytes()
    strips = [raw[r * 11 * 2:(r + 3) * 11 * 2] for r in range(0, 7, 3)]
    p = tmp_path / "p.tif"
    p.write_bytes(_tiff_bytes(11, 7, strips, 16, 1, 2, big_endian))
    np.testing.assert_array_equal(np.array(PIL.open(str(p))), a)            # libtiff agrees on the file
    np.testing.assert_array_equal(EI.read_tiff(str(p)), a.astype(np.float32) / np.float32(65535))
    # without compression the predictor tag is not applied
    q = tmp_path / "raw.tif"
    q.write_bytes(_tiff_bytes(11, 7, strips, 16, 1, 2, big_endian, compression=1))
    np.testing.assert_array_equal(EI.read_tiff(str(q)), d.astype(np.float32) / np.float32(65535))


def test_floating_point_predictor(tmp_path):
    rng = np.random.default_rng(6)
    a = rng.normal(0, 10, (5, 9)).astype(np.float32)
    rows = []
    for r in a:   # bytes in big-endian order, split into byte planes, differenced along the row
        b = np.frombuffer(r.astype(">f4").tobytes(), np.uint8).reshape(9, 4).T.ravel().astype(np.int64)
        b[1:] = (b[1:] - b[:-1]) % 256
        rows.append(b.astype(np.uint8).tobytes())
    p = tmp_path / "f.tif"
    p.write_bytes(_tiff_bytes(9, 5, [b"".join(rows)], 32, 3, 3, False))
    np.testing.assert_array_equal(np.array(PIL.open(str(p))), a)            # libtiff agrees on the file
    np.testing.assert_array_equal(EI.read_tiff(str(p)), a)


def test_tiff_writer_round_trip_and_pillow(tmp_path):
    a = np.random.default_rng(7).normal(0, 1, (33, 17)).astype(np.float32)
    p = str(tmp_path / "w.tif")
    EI.write_tiff(p, a)
    np.testing.assert_array_equal(EI.read_tiff(p), a)
    np.testing.assert_array_equal(np.array(PIL.open(p), np.float32), a)


def test_unsupported_tiff_samples_are_rejected(tmp_path):
    p = str(tmp_path / "i32.tif")
    PIL.fromarray(np.arange(12, dtype=np.int32).reshape(3, 4), mode="I").save(p)
    with pytest.raises(Unsupported):
        EI.read_tiff(p)
    q = tmp_path / "junk.tif"
    q.write_bytes(b"not a tiff at all")
    with pytest.raises(InvalidArgument):
        EI.read_tiff(str(q))


# ---- image-stack directories ------------------------------------------------

def test_image_stack_directory_round_trip(tmp_path):
    rng = np.random.default_rng(8)
    proj = rng.normal(0, 1, (5, 6, 7)).astype(np.float32)
    tilts = np.array([-30, -10.5, 0, 12.25, 44], np.float32)
    base = EI.ScannerGeometry.make((-3.5, -3, -20), (-3.5, -3, 20), (1, 0, 0), (0, 1, 0), (7, 6))
    EI.write_image_stack_directory(str(tmp_path), proj, tilts, base, (7, 6, 4))
    p, t, b = EI.open_projections(str(tmp_path))
    np.testing.assert_array_equal(p, proj)
    np.testing.assert_array_equal(t, tilts)
    for f in ("source", "detector", "horizontal_pitch", "vertical_pitch", "resolution"):
        assert list(getattr(b, f)) == list(getattr(base, f))


def _metadata(tmp_path, body):
    (tmp_path / "metadata.xml").write_text(
        '<?xml version="1.0"?>\n<geometry>\n <sourcePosition value="0 0 -5"/>\n <detectorPosition value="0 0 5"/>\n'
        ' <horizontalPitch value="1 0 0"/>\n <verticalPitch value="0 1 0"/>\n</geometry>\n' + body)


def test_logscale_and_projectionstack_entries(tmp_path):
    a = np.full((4, 4), 0.5, np.float32)
    EI.write_tiff(str(tmp_path / "a.tif"), a)
    pages = [np.full((4, 4), v, np.float32) for v in (1.0, 2.0)]
    PIL.fromarray(pages[0], mode="F").save(str(tmp_path / "s.tif"), save_all=True,
                                           append_images=[PIL.fromarray(pages[1], mode="F")])
    _metadata(tmp_path, '<projection><image filename="a.tif"/><tiltAngle value="3"/><logscale/></projection>\n'
                        '<projectionstack><image filename="s.tif"/><tiltAngle value="7"/></projectionstack>\n')
    # two stack pages share tilt angle 7: a two-dimensional stack
    with pytest.raises(Unsupported):
        EI.open_projections(str(tmp_path))
    _metadata(tmp_path, '<projection><image filename="a.tif"/><tiltAngle value="3"/><logscale/></projection>\n'
                        '<!-- comment --><projection><image filename="s.tif"/><tiltAngle value="-2"/></projection>\n')
    p, t, _ = EI.open_projections(str(tmp_path))
    np.testing.assert_array_equal(t, [3, -2])
    np.testing.assert_array_equal(p[0], np.exp(a))          # Image::convertToLinearScale
    np.testing.assert_array_equal(p[1], pages[0])


def test_malformed_metadata_names_the_line(tmp_path):
    (tmp_path / "metadata.xml").write_text('<geometry>\n<sourcePosition value="1 2 3">\n</geometry>\n')
    with pytest.raises(InvalidArgument, match=r"metadata.xml:3"):
        EI.open_projections(str(tmp_path))


# ---- MRC stacks ---------------------------------------------------------------

MODES = {0: np.uint8, 1: np.int16, 2: np.float32, 6: np.uint16}


def _write_mrc(path, data_zyx, mode, extended=0):
    nz, ny, nx = data_zyx.shape
    h = np.zeros(256, np.int32)
    h[:4] = (nx, ny, nz, mode)
    h[23] = extended
    raw = h.tobytes()
    raw = raw[:208] + b"MAP " + raw[212:]
    with open(path, "wb") as f:
        f.write(raw + bytes(range(256)) * (extended // 256) + b"\x07" * (extended % 256))
        f.write(np.ascontiguousarray(data_zyx, MODES[mode]).tobytes())


def _values(mode, shape, seed):
    rng = np.random.default_rng(seed)
    dt = MODES[mode]
    if mode == 2:
        return rng.normal(5, 3, shape).astype(np.float32)
    info = np.iinfo(dt)
    return rng.integers(max(info.min, -3000), min(info.max, 3000) + 1, shape).astype(dt)


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
@pytest.mark.parametrize("extended", [0, 300])
def test_mrc_stack_modes(tmp_path, mode, extended):
    data = _values(mode, (3, 5, 8), mode)
    p = str(tmp_path / "s.mrc")
    _write_mrc(p, data, mode, extended)
    proj, tilts, base = EI.open_projections(p)
    np.testing.assert_array_equal(proj, data.astype(np.float32))
    np.testing.assert_array_equal(tilts, np.zeros(3, np.float32))
    # MRCStack.cpp:250-262: parallel beams along -z through a detector of nx x ny unit pixels
    assert list(base.source) == [-4, -2.5, 8] and list(base.detector) == [-4, -2.5, -8]
    assert list(base.horizontal_pitch) == [1, 0, 0] and list(base.vertical_pitch) == [0, 1, 0]


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
def test_mrc_stack_log_transform_end_points(tmp_path, mode):
    data = _values(mode, (2, 6, 7), 10 + mode)
    p = str(tmp_path / "s.ali")
    _write_mrc(p, data, mode)
    proj, _, _ = EI.open_projections(p, logarithmize=True)
    f = data.astype(np.float64)
    lo, hi = f.min(), f.max()
    top = (hi - lo + 1) if mode in (1, 2) else (hi + 1)     # dataMax after MRCStack::init
    i_lo, i_hi = np.unravel_index(np.argmin(f), f.shape), np.unravel_index(np.argmax(f), f.shape)
    assert proj[i_hi] == 0.0
    if mode in (1, 2):
        np.testing.assert_allclose(proj[i_lo], np.log(top), rtol=1e-6)
    else:
        np.testing.assert_allclose(proj[i_lo], np.log(top / (lo + 1)), rtol=1e-6)
    # decreasing in the raw value
    order = np.argsort(f.ravel(), kind="stable")
    assert np.all(np.diff(proj.ravel()[order].astype(np.float64)) <= 1e-6)


def test_mrc_stack_tilt_file(tmp_path):
    p = str(tmp_path / "s.mrc")
    _write_mrc(p, _values(2, (4, 3, 3), 1), 2)
    (tmp_path / "s.tlt").write_text(" -60.5\n-20\n  10.25\n\n60\n")
    _, tilts, _ = EI.open_projections(p, str(tmp_path / "s.tlt"))
    np.testing.assert_array_equal(tilts, np.array([-60.5, -20, 10.25, 60], np.float32))
    (tmp_path / "bad.tlt").write_text("1\n2\n3\n")
    with pytest.raises(InvalidArgument, match="DIFFERS"):
        EI.open_projections(p, str(tmp_path / "bad.tlt"))


def test_mrc_stack_errors(tmp_path):
    p = str(tmp_path / "s.mrc")
    _write_mrc(p, np.zeros((2, 2, 2), np.float32), 2)
    raw = bytearray(open(p, "rb").read())
    raw[12:16] = struct.pack("<i", 3)   # mode 3: complex
    open(p, "wb").write(bytes(raw))
    with pytest.raises(Unsupported, match="mode 3"):
        EI.open_projections(p)
    q = str(tmp_path / "s.dat")
    _write_mrc(q, np.zeros((2, 2, 2), np.float32), 2)
    with pytest.raises(InvalidArgument, match="Unable to open image stack"):
        EI.open_projections(q)
    t = str(tmp_path / "t.mrc")
    open(t, "wb").write(open(q, "rb").read()[:1030])   # a valid header, data cut short
    with pytest.raises(InvalidArgument):
        EI.open_projections(t)


@needs_reference
def test_jason64_stack():
    d = os.path.join(REF_DATA, "jason_64_SART")
    proj, tilts, base = EI.open_projections(os.path.join(d, "jason64.ali"), os.path.join(d, "jason64.tlt"))
    raw = open(os.path.join(d, "jason64.ali"), "rb").read()
    assert proj.shape == (90, 64, 64)
    # mode 1 (int16) after an 11520-byte extended header
    ref = np.frombuffer(raw[1024 + 11520:], np.int16).reshape(90, 64, 64).astype(np.float32)
    np.testing.assert_array_equal(proj, ref)
    assert tilts[0] == np.float32(-62.78) and len(tilts) == 90
    assert list(base.source) == [-32, -32, 64]


# ---- MRC volumes ---------------------------------------------------------------

def test_mrc_writer_reproduces_a_reference_file_byte_for_byte(tmp_path):
    src = os.path.join(FILES, "voxelized_reference.mrc")
    vol = EI.read_mrc_volume(src, "xzy_order")
    out = str(tmp_path / "v.mrc")
    EI.write_mrc_volume(out, vol)          # defaults: float32, xzy_order
    assert open(out, "rb").read() == open(src, "rb").read()


ORDERS = {"xyz_order": (2, 1, 0), "xzy_order": (1, 2, 0), "yxz_order": (2, 0, 1), "yzx_order": (0, 2, 1),
          "zxy_order": (1, 0, 2), "zyx_order": (0, 1, 2)}   # volume axis (0 x, 1 y, 2 z) of sections, rows, columns


@pytest.mark.parametrize("order", sorted(ORDERS))
def test_mrc_orientations(tmp_path, order):
    vol = np.random.default_rng(9).normal(0, 1, (5, 6, 7)).astype(np.float32)   # z=5, y=6, x=7
    p = str(tmp_path / "o.mrc")
    EI.write_mrc_volume(p, vol, orientation=order)
    raw = open(p, "rb").read()
    h = np.frombuffer(raw[:1024], np.int32)
    s, r, c = ORDERS[order]
    res = (7, 6, 5)
    assert tuple(h[:4]) == (res[c], res[r], res[s], 2)
    file = np.frombuffer(raw[1024:], np.float32).reshape(res[s], res[r], res[c])
    np.testing.assert_array_equal(file, vol.transpose(2 - s, 2 - r, 2 - c))
    np.testing.assert_array_equal(EI.read_mrc_volume(p, order), vol)


def _header(p):
    raw = open(p, "rb").read(1024)
    return np.frombuffer(raw, np.int32), np.frombuffer(raw, np.float32)


@pytest.mark.parametrize("vt,mode,lo,hi,dt", [("unsigned8", 0, 0, 255, np.uint8), ("signed16", 1, -32768, 32767, np.int16),
                                             ("unsigned16", 6, 0, 65535, np.uint16)])
def test_integer_voxel_types_span_their_range(tmp_path, vt, mode, lo, hi, dt):
    vol = np.random.default_rng(10).normal(0, 1, (4, 5, 6)).astype(np.float32)
    for invert in (False, True):
        p = str(tmp_path / ("i%d.mrc" % invert))
        EI.write_mrc_volume(p, vol, voxel_type=vt, orientation="xyz_order", invert=invert)
        hi_, hf = _header(p)
        assert hi_[3] == mode
        data = np.frombuffer(open(p, "rb").read()[1024:], dt).reshape(vol.shape)
        a, b = (hi, lo) if invert else (lo, hi)
        assert data.ravel()[np.argmin(vol)] == a and data.ravel()[np.argmax(vol)] in (b, b - 1 if b > a else b + 1)
        assert hf[19] == a and hf[20] == b
        # monotone in the voxel value
        order = np.argsort(vol.ravel())
        steps = np.diff(data.ravel()[order].astype(np.int64))
        assert np.all(steps <= 0) if invert else np.all(steps >= 0)


def test_inverted_float_writes_unchanged_data(tmp_path):
    vol = np.random.default_rng(11).normal(2, 1, (3, 4, 5)).astype(np.float32)
    p = str(tmp_path / "f.mrc")
    EI.write_mrc_volume(p, vol, orientation="xyz_order", invert=True)
    _, hf = _header(p)
    # VolumeFileFormatWriter.cpp:24-29 writes float32 slice by slice from the
    # volume: only the header statistics are inverted
    np.testing.assert_array_equal(np.frombuffer(open(p, "rb").read()[1024:], np.float32).reshape(vol.shape), vol)
    # transformRange(min) = (min - max) + max in float arithmetic: within an ulp or two
    np.testing.assert_allclose([hf[19], hf[20]], [vol.max(), vol.min()], rtol=1e-6)


def test_header_statistics(tmp_path):
    vol = np.random.default_rng(12).normal(0, 1, (6, 6, 6)).astype(np.float32)
    p = str(tmp_path / "s.mrc")
    EI.write_mrc_volume(p, vol)
    _, hf = _header(p)
    assert hf[19] == vol.min() and hf[20] == vol.max()
    np.testing.assert_allclose(hf[21], vol.astype(np.float64).mean(), rtol=1e-5, atol=1e-6)
    # VolumeStatistics starts the maximum at the smallest positive normal float
    neg = -np.abs(vol) - 1
    EI.write_mrc_volume(p, neg)
    _, hf = _header(p)
    assert hf[20] == np.finfo(np.float32).tiny and hf[19] == neg.min()


@needs_reference
def test_reference_header_quirk_all_negative_volume(tmp_path):
    # ssd/SSDVolumeOriginal.mrc, written by the reference: maximum 1.1754944e-38
    src = os.path.join(REF_DATA, "ssd", "SSDVolumeOriginal.mrc")
    vol = EI.read_mrc_volume(src, "xzy_order")
    p = str(tmp_path / "q.mrc")
    EI.write_mrc_volume(p, vol)
    _, ours = _header(p)
    _, ref = _header(src)
    assert ours[19] == ref[19] and ours[20] == ref[20] == np.finfo(np.float32).tiny
    assert open(p, "rb").read()[1024:] == open(src, "rb").read()[1024:]
