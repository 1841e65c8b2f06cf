"""File formats and the command-line reconstruction of libetn.so.

Thin ctypes wrappers over the C ABI (include/etn/etn.h); every format is
implemented in the library (csrc/io.cpp, csrc/config.cpp, csrc/framework.cpp).
Volumes are numpy arrays [z][y][x]; projection stacks [count][nv][nu].
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from ._lib import ScannerGeometry, call, lib


def _b(s):
    return None if s is None else os.fspath(s).encode()


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def open_projections(path, tilt_angles=None, logarithmize=False):
    """ImageStackDatasourceFactory::instantiate: a directory with metadata.xml
    or an .mrc / .ali stack.  Returns (projections [count][nv][nu] float32,
    tilt angles (degrees), base ScannerGeometry)."""
    h = C.c_void_p()
    call("etn_projection_stack_open", _b(path), _b(tilt_angles) if tilt_angles else None, int(bool(logarithmize)),
         C.byref(h))
    try:
        nu, nv, n = C.c_uint32(), C.c_uint32(), C.c_uint32()
        base = ScannerGeometry()
        call("etn_projection_stack_info", h, C.byref(nu), C.byref(nv), C.byref(n), C.byref(base))
        proj = np.empty((n.value, nv.value, nu.value), np.float32)
        tilts = np.empty(n.value, np.float32)
        call("etn_projection_stack_read", h, proj.ctypes.data_as(C.c_void_p), tilts.ctypes.data_as(C.c_void_p))
        return proj, tilts, base
    finally:
        lib().etn_projection_stack_destroy(h)


def write_image_stack_directory(directory, projections, tilt_angles, base: ScannerGeometry, volume_resolution):
    """metadata.xml + projection_<i>.tif (float32) into an existing directory."""
    p = _f32(projections)
    t = _f32(tilt_angles)
    if p.ndim != 3 or len(t) != p.shape[0]:
        raise ValueError("projections must be [count][nv][nu] with one tilt angle each")
    res = np.asarray(volume_resolution, np.uint32)
    call("etn_image_stack_directory_write", _b(directory), p.ctypes.data_as(C.c_void_p), p.shape[0], p.shape[2],
         p.shape[1], t.ctypes.data_as(C.c_void_p), C.byref(base), res.ctypes.data_as(C.c_void_p))


def read_tiff(path, page=0):
    w, h = C.c_uint32(), C.c_uint32()
    call("etn_tiff_read", _b(path), int(page), C.byref(w), C.byref(h), None)
    out = np.empty((h.value, w.value), np.float32)
    call("etn_tiff_read", _b(path), int(page), C.byref(w), C.byref(h), out.ctypes.data_as(C.c_void_p))
    return out


def tiff_page_count(path):
    n = C.c_uint32()
    call("etn_tiff_page_count", _b(path), C.byref(n))
    return n.value


def write_tiff(path, image):
    a = _f32(image)
    if a.ndim != 2:
        raise ValueError("a TIFF image is 2-D")
    call("etn_tiff_write", _b(path), a.ctypes.data_as(C.c_void_p), a.shape[1], a.shape[0])


def write_mrc_volume(path, volume, voxel_type="float32", orientation="xzy_order", invert=False):
    """VolumeFileFormatWriter / MRCWriter; volume is [z][y][x]."""
    v = _f32(volume)
    if v.ndim != 3:
        raise ValueError("a volume is 3-D")
    res = np.array([v.shape[2], v.shape[1], v.shape[0]], np.uint32)
    call("etn_mrc_write_volume", _b(path), v.ctypes.data_as(C.c_void_p), res.ctypes.data_as(C.c_void_p),
         voxel_type.encode(), orientation.encode(), int(bool(invert)))


def read_mrc_volume(path, orientation="xzy_order"):
    """VolumeDeserializer::load with the file's axes mapped by `orientation`;
    returns [z][y][x] float32."""
    res = np.zeros(3, np.uint32)
    call("etn_mrc_read_volume", _b(path), orientation.encode(), res.ctypes.data_as(C.c_void_p), None)
    out = np.empty((int(res[2]), int(res[1]), int(res[0])), np.float32)
    call("etn_mrc_read_volume", _b(path), orientation.encode(), res.ctypes.data_as(C.c_void_p),
         out.ctypes.data_as(C.c_void_p))
    return out


def _argv(args):
    items = [b"etn"] + [os.fspath(a).encode() for a in args]
    arr = (C.c_char_p * (len(items) + 1))(*items, None)
    return len(items), arr


def resolve_config(args):
    """The parameter sets that `etn <args>` would run with, as a dict."""
    argc, argv = _argv(args)
    n = C.c_size_t()
    call("etn_config_resolve", argc, argv, None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    call("etn_config_resolve", argc, argv, buf, n.value + 1, C.byref(n))
    return json.loads(buf.value.decode())


def main(args):
    """Runs the command-line reconstruction in this process; returns its exit code."""
    argc, argv = _argv(args)
    return int(lib().etn_main(argc, argv))
