"""libetn.so on the CPU: it loads, exports every entry point include/etn/etn.h
declares, and its host-side geometry / projection orders / error mapping
agree with the oracle (no device work)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import _lib
from ettention_amd import api as A
from oracle import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "etn", "etn.h")


def declared_entry_points():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(etn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_entry_points()
    assert len(names) >= 25
    L = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes signature table covers exactly the header
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """The ctypes mirrors in _lib.py have the C compiler's field offsets and sizes."""
    import subprocess
    exe = str(tmp_path / "desc_layout")
    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run(["gcc", "-std=c99", "-I" + os.path.dirname(os.path.dirname(HEADER)),
                    os.path.join(here, "native", "desc_layout.c"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    mirrors = {"etn_reconstruction_desc": _lib.ReconstructionDesc, "etn_scanner_geometry": _lib.ScannerGeometry,
               "etn_volume_geometry": _lib.VolumeGeometry}
    checked = 0
    for line in out:
        if not line:
            continue
        struct, field, value = line.split()
        cls = mirrors[struct]
        if field == "sizeof":
            assert C.sizeof(cls) == int(value), struct
        else:
            name = "lambda_" if field == "lambda" else field
            assert getattr(cls, name).offset == int(value), (struct, field)
        checked += 1
    assert checked == 24


def _to_oracle(g):
    return O.Scanner.make(g.source, g.detector, g.horizontal_pitch, g.vertical_pitch, g.resolution)


BASES = [
    E.ScannerGeometry.mrc_base(512, 512),
    E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256)),
    E.ScannerGeometry.make((-16, -16, -256), (-16, -16, 256), (0.125, 0, 0), (0, 0.125, 0), (256, 256)),
]


@pytest.mark.parametrize("base", BASES)
def test_host_geometry_matches_oracle_bit_for_bit(base):
    ob = _to_oracle(base)
    for t in (-60.0, -37.5, -1.0, 0.0, 13.0, 45.0, 59.0, 80.0, 90.0, 225.0):
        for xt in (0.0, 3.5):
            g = A.rotate(base, t, xt)
            og = O.rotate(ob, t, xt)
            assert list(g.source) == list(og.source)
            assert list(g.detector) == list(og.detector)
            assert list(g.horizontal_pitch) == list(og.hpitch)
            assert list(g.vertical_pitch) == list(og.vpitch)
            np.testing.assert_array_equal(A.ray_direction(g), O.ray_direction(og))
            np.testing.assert_array_equal(A.projection_matrix(g), O.projection_matrix(og))
            for n in (1, 4, 9, 16):
                np.testing.assert_array_equal(A.bp_sample_pattern(g, (1, 1, 1), n),
                                              O.bp_sample_pattern(og, (1, 1, 1), n))
            assert A.traversal_length_through_bbox(g, (-8, -8, -4), (8, 8, 4)) == \
                O.traversal_length_bbox(og, (-8, -8, -4), (8, 8, 4))


def test_subvolumes_match_oracle():
    for res, S in [((64, 64, 64), 4), ((8, 6, 10), 4), ((512, 512, 256), 3), ((16, 16, 7), 7)]:
        v = E.VolumeGeometry.make(res, subvolumes=S)
        ov = O.Volume.make(res, slabs=S)
        for i in range(S):
            a = A.subvolume(v, i)
            b = O.slab(ov, i)
            np.testing.assert_array_equal(a[0], b[0])
            assert a[1] == b[1]
            np.testing.assert_array_equal(a[2], b[2])
            np.testing.assert_array_equal(a[3], b[3])


def test_projection_orders_match_oracle():
    for scheme in ("identity", "random", "maxangle", "buckets"):
        for n in (1, 2, 5, 6, 61, 121, 180):
            for skip in ((), (0,), (1, 3)):
                if len(skip) >= n:
                    continue
                np.testing.assert_array_equal(A.projection_order(scheme, n, skip),
                                              O.projection_order(scheme, n, skip))


def test_errors_map_to_status_codes():
    g = A.rotate(BASES[1], 0.0)
    with pytest.raises(E.InvalidArgument, match="non-square"):
        A.bp_sample_pattern(g, (1, 1, 1), 3)
    with pytest.raises(E.InvalidArgument, match="zero samples"):
        A.bp_sample_pattern(g, (1, 1, 1), 0)
    with pytest.raises(E.InvalidArgument, match="Illegal ProjectionSetType"):
        A.projection_order("sideways", 4)
    with pytest.raises(E.InvalidArgument):
        A.subvolume(E.VolumeGeometry.make((4, 4, 3), subvolumes=4), 0)
    with pytest.raises(E.InvalidArgument, match="Length of vector is 0"):
        A.ray_direction(E.ScannerGeometry.make((0, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (4, 4)))


def _desc(**kw):
    d = _lib.ReconstructionDesc()
    d.algorithm = kw.get("algorithm", b"sart")
    d.block_size = kw.get("block_size", 0)
    d.iterations = 1
    d.lambda_ = 0.3
    d.volume = E.VolumeGeometry.make((8, 8, 8))
    d.base_geometry = E.ScannerGeometry.mrc_base(8, 8)
    tilts = kw.get("tilts", np.zeros(3, np.float32))
    d.projection_count = len(tilts)
    d.tilt_angles = tilts.ctypes.data_as(C.POINTER(C.c_float)) if len(tilts) else None
    d.projections = C.c_void_p(1)
    d.projection_iteration = kw.get("order", b"identity")
    return d, tilts


@pytest.mark.parametrize("kw,exc,msg", [
    (dict(algorithm=b"blockiterative", block_size=0), E.InvalidArgument, "Illegal ProjectionBlockSize"),
    (dict(algorithm=b"art"), E.Unsupported, "unknown reconstruction algorithm"),
    (dict(tilts=np.zeros(0, np.float32)), E.InvalidArgument, "No projections given"),
    (dict(order=b"zigzag"), E.InvalidArgument, "Illegal ProjectionSetType"),
])
def test_reconstruction_preconditions_fail_before_device_work(kw, exc, msg):
    # BlockIterativeProjectionAccess.cpp:16-19,25-40 and AlgorithmParameterSet: checked on the
    # host before any device allocation.
    d, _keep = _desc(**kw)
    h = C.c_void_p()
    st = _lib.lib().etn_reconstruction_create(C.byref(d), None, C.byref(h))
    with pytest.raises(exc, match=msg):
        _lib.check(st)
    assert not h.value
