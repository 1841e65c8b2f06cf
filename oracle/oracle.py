"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path (libetn.so and
the ettention_amd package) never imports this module.

Each wrapper forwards to the C function of the same name in etn_oracle.c,
which cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

VOXEL_F32, VOXEL_F16, VOXEL_BF16 = 0, 1, 2


class Scanner(C.Structure):
    _fields_ = [("source", C.c_float * 3), ("detector", C.c_float * 3),
                ("hpitch", C.c_float * 3), ("vpitch", C.c_float * 3),
                ("res", C.c_uint32 * 2)]

    @staticmethod
    def make(source, detector, hpitch, vpitch, res):
        s = Scanner()
        s.source[:] = [float(x) for x in source]
        s.detector[:] = [float(x) for x in detector]
        s.hpitch[:] = [float(x) for x in hpitch]
        s.vpitch[:] = [float(x) for x in vpitch]
        s.res[:] = [int(x) for x in res]
        return s

    def as_tuple(self):
        return (list(self.source), list(self.detector), list(self.hpitch), list(self.vpitch), list(self.res))


class Volume(C.Structure):
    _fields_ = [("res", C.c_uint32 * 3), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
                ("slabs", C.c_uint32)]

    @staticmethod
    def make(res, bmin=None, bmax=None, slabs=1):
        v = Volume()
        v.res[:] = [int(x) for x in res]
        if bmin is None:
            bmin = [-0.5 * r for r in res]
            bmax = [0.5 * r for r in res]
        v.bmin[:] = [float(x) for x in bmin]
        v.bmax[:] = [float(x) for x in bmax]
        v.slabs = int(slabs)
        return v


class Recon(C.Structure):
    _fields_ = [("vol", Volume), ("voxel_type", C.c_int), ("base", Scanner), ("nproj", C.c_uint32),
                ("tilt_deg", C.POINTER(C.c_float)), ("xtilt_deg", C.c_float),
                ("projections", C.POINTER(C.c_float)), ("block_size", C.c_uint32),
                ("order", C.c_char_p), ("skip", C.POINTER(C.c_uint32)), ("nskip", C.c_uint32),
                ("lambda_", C.c_float), ("use_loc", C.c_int), ("fp_samples", C.c_uint32),
                ("bp_samples", C.c_uint32), ("mask", C.POINTER(C.c_uint8))]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        fp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        S, V = C.POINTER(Scanner), C.POINTER(Volume)
        L.orc_rotate.argtypes = [S, f32, f32, S]
        L.orc_rotate_f64.argtypes = [S, C.c_double, C.c_double, fp]
        L.orc_ray_direction.argtypes = [S, fp]
        L.orc_projection_matrix.argtypes = [S, fp]
        L.orc_projection_matrix.restype = i32
        L.orc_traversal_length_bbox.argtypes = [S, fp, fp]
        L.orc_traversal_length_bbox.restype = f32
        L.orc_bp_sample_pattern.argtypes = [S, fp, u32, fp]
        L.orc_bp_sample_pattern.restype = i32
        L.orc_voxel_size.argtypes = [V, fp]
        L.orc_slab.argtypes = [V, u32, fp, fp, fp, fp]
        L.orc_slab.restype = i32
        L.orc_fp.argtypes = [V, i32, fp, fp, u32, S, u32, fp, fp]
        L.orc_fp.restype = i32
        L.orc_compare.argtypes = [C.c_size_t, fp, fp, fp]
        L.orc_long_object_compensate.argtypes = [C.c_size_t, fp, fp, f32]
        L.orc_ray_length.argtypes = [V, S, fp]
        L.orc_bp.argtypes = [V, i32, fp, fp, u32, S, u32, fp, fp, f32, u32]
        L.orc_bp.restype = i32
        L.orc_fp_f64.argtypes = [V, fp, u32, fp, fp, u32, fp, fp]
        L.orc_fp_f64.restype = i32
        L.orc_bp_f64.argtypes = [V, fp, u32, fp, fp, u32, fp, fp, C.c_double, u32]
        L.orc_bp_f64.restype = i32
        L.orc_siddon.argtypes = [V, fp, fp, fp, fp]
        L.orc_siddon.restype = C.c_double
        for name in ("orc_f32_to_f16_rte", "orc_f32_to_f16_trunc", "orc_f32_to_bf16_rne"):
            getattr(L, name).argtypes = [f32]
            getattr(L, name).restype = C.c_uint16
        for name in ("orc_f16_to_f32", "orc_bf16_to_f32"):
            getattr(L, name).argtypes = [C.c_uint16]
            getattr(L, name).restype = f32
        L.orc_mt19937_nth.argtypes = [u32, u32]
        L.orc_mt19937_nth.restype = u32
        L.orc_mt19937_state_words.restype = u32
        L.orc_mt19937_seed.argtypes = [fp, u32]
        L.orc_mt19937_next.argtypes = [fp]
        L.orc_mt19937_next.restype = u32
        L.orc_boost_uniform_int.argtypes = [fp, u32, u32]
        L.orc_boost_uniform_int.restype = u32
        L.orc_projection_order.argtypes = [C.c_char_p, u32, fp, u32, fp, fp]
        L.orc_projection_order.restype = i32
        L.orc_block_iterative.argtypes = [C.POINTER(Recon), fp, u32]
        L.orc_block_iterative.restype = i32
        L.orc_block_iterative_f64.argtypes = [C.POINTER(Recon), fp, u32]
        L.orc_block_iterative_f64.restype = i32
        L.orc_residual_norm.argtypes = [C.POINTER(Recon), fp]
        L.orc_residual_norm.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def rotate(base: Scanner, tilt_deg, xtilt_deg=0.0) -> Scanner:
    out = Scanner()
    lib().orc_rotate(C.byref(base), float(tilt_deg), float(xtilt_deg), C.byref(out))
    return out


def rotate_f64(base: Scanner, tilt_deg, xtilt_deg=0.0):
    out = np.zeros(12, np.float64)
    lib().orc_rotate_f64(C.byref(base), float(tilt_deg), float(xtilt_deg), _p(out))
    return out


def ray_direction(g: Scanner):
    out = np.zeros(3, np.float32)
    lib().orc_ray_direction(C.byref(g), _p(out))
    return out


def projection_matrix(g: Scanner):
    out = np.zeros(16, np.float32)
    if lib().orc_projection_matrix(C.byref(g), _p(out)):
        raise ValueError("matrix not invertible")
    return out


def traversal_length_bbox(g: Scanner, bmin, bmax):
    a = np.asarray(bmin, np.float32)
    b = np.asarray(bmax, np.float32)
    return float(lib().orc_traversal_length_bbox(C.byref(g), _p(a), _p(b)))


def bp_sample_pattern(g: Scanner, voxel_size, samples):
    vs = np.asarray(voxel_size, np.float32)
    out = np.zeros(4 * max(samples, 1), np.float32)
    if lib().orc_bp_sample_pattern(C.byref(g), _p(vs), int(samples), _p(out)):
        raise ValueError("samples must be a non-zero perfect square")
    return out.reshape(-1, 4)


def slab(v: Volume, index):
    sres = np.zeros(3, np.uint32)
    z0 = np.zeros(1, np.uint32)
    smin = np.zeros(3, np.float32)
    smax = np.zeros(3, np.float32)
    if lib().orc_slab(C.byref(v), int(index), _p(sres), _p(z0), _p(smin), _p(smax)):
        raise ValueError("bad slab")
    return sres, int(z0[0]), smin, smax


def forward_project(v: Volume, volume, g: Scanner, samples=1, slab=0, mask=None,
                    voxel_type=VOXEL_F32, proj=None, length=None):
    nu, nv = g.res[0], g.res[1]
    if proj is None:
        proj = np.zeros((nv, nu), np.float32)
    if length is None:
        length = np.zeros((nv, nu), np.float32)
    vol = np.ascontiguousarray(volume)
    mk = None if mask is None else _p(np.ascontiguousarray(mask, np.uint8))
    if lib().orc_fp(C.byref(v), voxel_type, _p(vol), mk, int(slab), C.byref(g), int(samples),
                    _p(proj), _p(length)):
        raise ValueError("bad slab")
    return proj, length


def compare(real, virt):
    res = np.empty_like(real, dtype=np.float32)
    lib().orc_compare(real.size, _p(np.ascontiguousarray(real, np.float32)),
                      _p(np.ascontiguousarray(virt, np.float32)), _p(res))
    return res


def long_object_compensate(proj, length, c):
    lib().orc_long_object_compensate(proj.size, _p(proj), _p(length), float(c))


def ray_length(v: Volume, g: Scanner):
    out = np.zeros((g.res[1], g.res[0]), np.float32)
    lib().orc_ray_length(C.byref(v), C.byref(g), _p(out))
    return out


def back_project(v: Volume, volume, g: Scanner, residual, raylen, lam=1.0, block_size=1,
                 samples=1, slab=0, mask=None, voxel_type=VOXEL_F32):
    mk = None if mask is None else _p(np.ascontiguousarray(mask, np.uint8))
    if lib().orc_bp(C.byref(v), voxel_type, _p(volume), mk, int(slab), C.byref(g), int(samples),
                    _p(np.ascontiguousarray(residual, np.float32)),
                    _p(np.ascontiguousarray(raylen, np.float32)), float(lam), int(block_size)):
        raise ValueError("bad back projection arguments")
    return volume


def forward_project_f64(v: Volume, volume, geom12, det_res, samples=1, slab=0, proj=None, length=None):
    nu, nv = det_res
    if proj is None:
        proj = np.zeros((nv, nu), np.float64)
    if length is None:
        length = np.zeros((nv, nu), np.float64)
    dr = np.asarray(det_res, np.uint32)
    lib().orc_fp_f64(C.byref(v), _p(np.ascontiguousarray(volume, np.float64)), int(slab),
                     _p(np.asarray(geom12, np.float64)), _p(dr), int(samples), _p(proj), _p(length))
    return proj, length


def back_project_f64(v: Volume, volume, geom12, det_res, residual, raylen, lam=1.0, block_size=1,
                     samples=1, slab=0):
    dr = np.asarray(det_res, np.uint32)
    if lib().orc_bp_f64(C.byref(v), _p(volume), int(slab), _p(np.asarray(geom12, np.float64)), _p(dr),
                        int(samples), _p(np.ascontiguousarray(residual, np.float64)),
                        _p(np.ascontiguousarray(raylen, np.float64)), float(lam), int(block_size)):
        raise ValueError("bad back projection arguments")
    return volume


def siddon(v: Volume, volume, origin, direction):
    chord = np.zeros(1, np.float64)
    val = lib().orc_siddon(C.byref(v), _p(np.ascontiguousarray(volume, np.float64)),
                           _p(np.asarray(origin, np.float64)), _p(np.asarray(direction, np.float64)),
                           _p(chord))
    return val, float(chord[0])


def projection_order(kind, count, skip=()):
    out = np.zeros(max(count, 1), np.uint32)
    n = np.zeros(1, np.uint32)
    sk = np.asarray(list(skip) or [0], np.uint32)
    if lib().orc_projection_order(kind.encode(), int(count), _p(sk), len(skip), _p(out), _p(n)):
        raise ValueError("unknown projection iteration scheme: %s" % kind)
    return out[: int(n[0])].copy()


def num_threads():
    return int(lib().orc_num_threads())


def mt19937_nth(seed, n):
    return int(lib().orc_mt19937_nth(seed, n))


class Mt19937:
    """The oracle's mt19937 engine and its boost::uniform_int<> mapping
    (the draws of ProjectionSetRandom::reorder)."""

    def __init__(self, seed=5489):
        self.state = np.zeros(int(lib().orc_mt19937_state_words()), np.uint32)
        lib().orc_mt19937_seed(_p(self.state), seed)

    def next(self):
        return int(lib().orc_mt19937_next(_p(self.state)))

    def uniform_int(self, lo, hi):
        return int(lib().orc_boost_uniform_int(_p(self.state), lo, hi))


class ReconSpec:
    """Inputs of one block-iterative reconstruction (mirrors orc_recon)."""

    def __init__(self, vol: Volume, base: Scanner, tilts, projections, block_size=1, order="identity",
                 lam=0.3, use_loc=False, fp_samples=1, bp_samples=1, skip=(), xtilt=0.0,
                 voxel_type=VOXEL_F32, mask=None):
        self.tilts = np.ascontiguousarray(tilts, np.float32)
        self.projections = np.ascontiguousarray(projections, np.float32)
        self.skip = np.asarray(list(skip) or [0], np.uint32)
        self.order = order.encode()
        self.mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        r = Recon()
        r.vol = vol
        r.voxel_type = voxel_type
        r.base = base
        r.nproj = len(self.tilts)
        r.tilt_deg = self.tilts.ctypes.data_as(C.POINTER(C.c_float))
        r.xtilt_deg = float(xtilt)
        r.projections = self.projections.ctypes.data_as(C.POINTER(C.c_float))
        r.block_size = int(block_size)
        r.order = self.order
        r.skip = self.skip.ctypes.data_as(C.POINTER(C.c_uint32))
        r.nskip = len(skip)
        r.lambda_ = float(lam)
        r.use_loc = int(bool(use_loc))
        r.fp_samples = int(fp_samples)
        r.bp_samples = int(bp_samples)
        r.mask = None if self.mask is None else self.mask.ctypes.data_as(C.POINTER(C.c_uint8))
        self.r = r

    def run(self, volume, iterations):
        if lib().orc_block_iterative(C.byref(self.r), _p(volume), int(iterations)):
            raise ValueError("reconstruction failed")
        return volume

    def run_f64(self, volume, iterations):
        if lib().orc_block_iterative_f64(C.byref(self.r), _p(volume), int(iterations)):
            raise ValueError("reconstruction failed")
        return volume

    def residual_norm(self, volume):
        return float(lib().orc_residual_norm(C.byref(self.r), _p(volume)))


def mrc_base_scanner(nx, ny):
    """MRCStack base geometry: source (-nx/2, -ny/2, nx), detector (-nx/2, -ny/2, -nx),
    unit pitches (ettention/libReconstruction/src/io/datasource/MRCStack.cpp:250-262)."""
    return Scanner.make((-0.5 * nx, -0.5 * ny, float(nx)), (-0.5 * nx, -0.5 * ny, -float(nx)),
                        (1, 0, 0), (0, 1, 0), (nx, ny))


def test_base_scanner():
    """ProjectionTestBase geometry: 256x256 detector, pitch 0.5, source (-64,-64,-128),
    detector (-64,-64,128) (ettention/libReconstruction/src/framework/test/ProjectionTestBase.cpp:105-119)."""
    return Scanner.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))
