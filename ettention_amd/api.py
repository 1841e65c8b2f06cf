"""Python front end of libetn.so for tests and benchmarks.

Device memory and streams come from PyTorch-ROCm tensors (plumbing only); all
computation happens in libetn.so.  Names follow the reference operators:
forward_project ~ ForwardProjectionOperator::Execute, back_project ~
BackProjectionOperator::execute, BlockIterativeReconstructionOperator ~ the
C++ class of the same name.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (ETN_RECON_FLAG_16BIT_PROJECTIONS, ETN_ALL_SUBVOLUMES, ETN_VOXEL_BFLOAT16, ETN_VOXEL_FLOAT16, ETN_VOXEL_FLOAT32,
                   ReconstructionDesc, ScannerGeometry, VolumeGeometry, call, lib)

_VOXEL_TYPES = {"float": ETN_VOXEL_FLOAT32, "float32": ETN_VOXEL_FLOAT32, "half": ETN_VOXEL_FLOAT16,
                "float16": ETN_VOXEL_FLOAT16, "bfloat16": ETN_VOXEL_BFLOAT16, "bf16": ETN_VOXEL_BFLOAT16}


def voxel_type_id(vt):
    if isinstance(vt, int):
        return vt
    return _VOXEL_TYPES[vt]


def _torch():
    import torch
    return torch


def _stream(stream=None):
    if stream is not None:
        return C.c_void_p(stream)
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _check(t, name, dtypes, shape):
    """Rejects a device tensor the C side would read past or misinterpret:
    the library reads exactly prod(shape) elements of the given type."""
    torch = _torch()
    if t is None:
        return
    dts = tuple(getattr(torch, d) for d in dtypes)
    if t.dtype not in dts:
        raise ValueError("%s must be %s, got %s" % (name, " or ".join(dtypes), t.dtype))
    if tuple(t.shape) != tuple(int(x) for x in shape):
        raise ValueError("%s must have shape %s, got %s" % (name, tuple(int(x) for x in shape), tuple(t.shape)))


def _volume_shape(vol_geom):
    return (vol_geom.resolution[2], vol_geom.resolution[1], vol_geom.resolution[0])


def _image_shape(g):
    return (g.resolution[1], g.resolution[0])


def _voxel_type_of(t):
    torch = _torch()
    return {torch.float32: ETN_VOXEL_FLOAT32, torch.float16: ETN_VOXEL_FLOAT16,
            torch.bfloat16: ETN_VOXEL_BFLOAT16}[t.dtype]


# ---- host geometry ----------------------------------------------------------

def rotate(base: ScannerGeometry, tilt_deg, x_axis_tilt_deg=0.0) -> ScannerGeometry:
    out = ScannerGeometry()
    call("etn_rotate_scanner_geometry", C.byref(base), float(tilt_deg), float(x_axis_tilt_deg), C.byref(out))
    return out


def ray_direction(g: ScannerGeometry):
    out = np.zeros(3, np.float32)
    call("etn_ray_direction", C.byref(g), out.ctypes.data_as(C.c_void_p))
    return out


def projection_matrix(g: ScannerGeometry):
    out = np.zeros(16, np.float32)
    call("etn_projection_matrix", C.byref(g), out.ctypes.data_as(C.c_void_p))
    return out


def traversal_length_through_bbox(g: ScannerGeometry, bmin, bmax):
    a = np.asarray(bmin, np.float32)
    b = np.asarray(bmax, np.float32)
    out = np.zeros(1, np.float32)
    call("etn_traversal_length_through_bbox", C.byref(g), a.ctypes.data_as(C.c_void_p),
         b.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return float(out[0])


def bp_sample_pattern(g: ScannerGeometry, voxel_size, samples):
    vs = np.asarray(voxel_size, np.float32)
    out = np.zeros(4 * max(int(samples), 1), np.float32)
    call("etn_bp_sample_pattern", C.byref(g), vs.ctypes.data_as(C.c_void_p), int(samples),
         out.ctypes.data_as(C.c_void_p))
    return out.reshape(-1, 4)


def subvolume(v: VolumeGeometry, index):
    res = np.zeros(3, np.uint32)
    first = np.zeros(1, np.uint32)
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    call("etn_subvolume", C.byref(v), int(index), res.ctypes.data_as(C.c_void_p), first.ctypes.data_as(C.c_void_p),
         bmin.ctypes.data_as(C.c_void_p), bmax.ctypes.data_as(C.c_void_p))
    return res, int(first[0]), bmin, bmax


def projection_order(scheme, count, skip=()):
    out = np.zeros(max(int(count), 1), np.uint32)
    n = np.zeros(1, np.uint32)
    sk = np.asarray(list(skip) or [0], np.uint32)
    call("etn_projection_order", scheme.encode(), int(count), sk.ctypes.data_as(C.c_void_p), len(skip),
         out.ctypes.data_as(C.c_void_p), n.ctypes.data_as(C.c_void_p))
    return out[: int(n[0])].copy()


# ---- kernels ----------------------------------------------------------------

def forward_project(vol_geom: VolumeGeometry, volume, g: ScannerGeometry, projection, traversal_length,
                    samples=1, subvolume=0, mask=None, stream=None):
    """Adds the subvolume's line integrals into projection / traversal_length
    (float32 device tensors of shape (nv, nu))."""
    _check(volume, "volume", ("float32", "float16", "bfloat16"), _volume_shape(vol_geom))
    _check(mask, "mask", ("uint8",), _volume_shape(vol_geom))
    _check(projection, "projection", ("float32",), _image_shape(g))
    _check(traversal_length, "traversal_length", ("float32",), _image_shape(g))
    call("etn_forward_project", C.byref(vol_geom), _voxel_type_of(volume), _ptr(volume), _ptr(mask),
         int(subvolume), C.byref(g), int(samples), _ptr(projection), _ptr(traversal_length), _stream(stream))


def back_project(vol_geom: VolumeGeometry, volume, g: ScannerGeometry, residual, ray_length, lam=1.0,
                 block_size=1, samples=1, subvolume=ETN_ALL_SUBVOLUMES, mask=None, stream=None):
    _check(volume, "volume", ("float32", "float16", "bfloat16"), _volume_shape(vol_geom))
    _check(mask, "mask", ("uint8",), _volume_shape(vol_geom))
    _check(residual, "residual", ("float32",), _image_shape(g))
    _check(ray_length, "ray_length", ("float32",), _image_shape(g))
    call("etn_back_project", C.byref(vol_geom), _voxel_type_of(volume), _ptr(volume), _ptr(mask),
         int(subvolume), C.byref(g), int(samples), _ptr(residual), _ptr(ray_length), float(lam),
         int(block_size), _stream(stream))


def compare(real, virtual, residual, stream=None):
    nv, nu = real.shape[-2], real.shape[-1]
    for name, t in (("real", real), ("virtual", virtual), ("residual", residual)):
        _check(t, name, ("float32",), (nv, nu))
    call("etn_compare", int(nu), int(nv), _ptr(real), _ptr(virtual), _ptr(residual), _stream(stream))


def long_object_compensate(projection, traversal_length, constant_length, stream=None):
    nv, nu = projection.shape[-2], projection.shape[-1]
    _check(projection, "projection", ("float32",), (nv, nu))
    _check(traversal_length, "traversal_length", ("float32",), (nv, nu))
    call("etn_long_object_compensate", int(nu), int(nv), _ptr(projection), _ptr(traversal_length),
         float(constant_length), _stream(stream))


def compute_ray_length(vol_geom: VolumeGeometry, g: ScannerGeometry, out, stream=None):
    _check(out, "out", ("float32",), _image_shape(g))
    call("etn_compute_ray_length", C.byref(vol_geom), C.byref(g), _ptr(out), _stream(stream))


def ramp_filter(projection, out, stream=None):
    """ConvolutionOperator with WBP's 128-tap RamLak kernel (etn_ramp_filter);
    float32 device tensors of shape (nv, nu)."""
    nv, nu = projection.shape[-2], projection.shape[-1]
    _check(projection, "projection", ("float32",), (nv, nu))
    _check(out, "out", ("float32",), (nv, nu))
    call("etn_ramp_filter", int(nu), int(nv), _ptr(projection), _ptr(out), _stream(stream))


def sum_squared_difference(a, b, out, stream=None):
    _check(a, "a", ("float32",), tuple(a.shape))
    _check(b, "b", ("float32",), tuple(a.shape))
    _check(out, "out", ("float64",), tuple(out.shape))
    if out.numel() < 1:
        raise ValueError("out must hold one double")
    call("etn_sum_squared_difference_f64", _ptr(a), _ptr(b), a.numel(), _ptr(out), _stream(stream))


def sum_squared_difference_noalloc(a, b, out, stream=None):
    """sum_squared_difference without device allocation: the partial sums use
    the library's static per-GPU scratch, so overlapping calls on one GPU race."""
    _check(a, "a", ("float32",), tuple(a.shape))
    if b is not None:
        _check(b, "b", ("float32",), tuple(a.shape))
    _check(out, "out", ("float64",), tuple(out.shape))
    if out.numel() < 1:
        raise ValueError("out must hold one double")
    call("etn_sum_squared_difference_f64_noalloc", _ptr(a), _ptr(b), a.numel(), _ptr(out), _stream(stream))


def convert_volume(src, dst, stream=None):
    if dst.numel() != src.numel():
        raise ValueError("dst must hold as many voxels as src (%d), got %d" % (src.numel(), dst.numel()))
    call("etn_convert_volume", _ptr(src), _voxel_type_of(src), _ptr(dst), _voxel_type_of(dst), src.numel(),
         _stream(stream))


# ---- reconstruction ---------------------------------------------------------

def _make_desc(*, volume, base_geometry, tilt_angles, projections, algorithm="sart", block_size=0, iterations=1,
               lam=0.3, use_long_object_compensation=False, fp_samples=1, bp_samples=1,
               projection_iteration="random", skip_projections=(), voxel_type="float", init_value=0.0,
               x_axis_tilt=0.0, mask=None, flags=0, geometries=None):
    """etn_reconstruction_desc plus the host objects it points into (keep them alive).
    geometries: optional sequence of ScannerGeometry, one per projection, used
    instead of rotating base_geometry by the tilt angles."""
    tilts = np.ascontiguousarray(tilt_angles, np.float32)
    # the library reads len(tilts) * nv * nu values of projections and
    # nx * ny * nz bytes of mask.  16-bit projections (stored in the volume's
    # 16-bit type) set ETN_RECON_FLAG_16BIT_PROJECTIONS.
    torch = _torch()
    vt = voxel_type_id(voxel_type)
    sixteen = {ETN_VOXEL_FLOAT16: "float16", ETN_VOXEL_BFLOAT16: "bfloat16"}
    if projections is not None and projections.dtype in (torch.float16, torch.bfloat16):
        if vt not in sixteen:
            raise ValueError("16-bit projections need voxel_type 'half' or 'bfloat16' (they take the volume's type)")
        _check(projections, "projections", (sixteen[vt],), (len(tilts),) + _image_shape(base_geometry))
        flags = int(flags) | ETN_RECON_FLAG_16BIT_PROJECTIONS
    else:
        _check(projections, "projections", ("float32",), (len(tilts),) + _image_shape(base_geometry))
    _check(mask, "mask", ("uint8",), _volume_shape(volume))
    skip = np.asarray(list(skip_projections) or [0], np.uint32)
    alg = algorithm.encode()
    order = projection_iteration.encode() if projection_iteration else None
    d = ReconstructionDesc()
    d.algorithm = alg
    d.block_size = int(block_size)
    d.iterations = int(iterations)
    d.lambda_ = float(lam)
    d.use_long_object_compensation = int(bool(use_long_object_compensation))
    d.fp_samples = int(fp_samples)
    d.bp_samples = int(bp_samples)
    d.projection_iteration = order
    d.skip_projections = skip.ctypes.data_as(C.POINTER(C.c_uint32))
    d.skip_count = len(skip_projections)
    d.voxel_type = voxel_type_id(voxel_type)
    d.volume = volume
    d.init_value = float(init_value)
    d.base_geometry = base_geometry
    d.projection_count = len(tilts)
    d.tilt_angles = tilts.ctypes.data_as(C.POINTER(C.c_float))
    d.x_axis_tilt = float(x_axis_tilt)
    d.projections = _ptr(projections)
    d.mask = _ptr(mask)
    d.flags = int(flags)
    geoms = None
    if geometries is not None:
        if len(geometries) != len(tilts):
            raise ValueError("geometries must hold one geometry per projection (%d), got %d"
                             % (len(tilts), len(geometries)))
        geoms = (ScannerGeometry * len(geometries))(*geometries)
        d.geometries = C.cast(geoms, C.POINTER(ScannerGeometry))
    return d, (tilts, skip, alg, order, projections, mask, geoms)


class BlockIterativeReconstructionOperator:
    """Wraps etn_reconstruction_* (BlockIterativeReconstructionOperator).

    projections: float32 device tensor (count, nv, nu) in stack order;
    tilt_angles: host sequence of degrees."""

    def __init__(self, *, volume: VolumeGeometry, base_geometry: ScannerGeometry, tilt_angles, projections,
                 algorithm="sart", block_size=0, iterations=1, lam=0.3, use_long_object_compensation=False,
                 fp_samples=1, bp_samples=1, projection_iteration="random", skip_projections=(),
                 voxel_type="float", init_value=0.0, x_axis_tilt=0.0, mask=None, flags=0, stream=None,
                 geometries=None):
        d, self._keep = _make_desc(volume=volume, base_geometry=base_geometry, tilt_angles=tilt_angles,
                                   projections=projections, algorithm=algorithm, block_size=block_size,
                                   iterations=iterations, lam=lam,
                                   use_long_object_compensation=use_long_object_compensation,
                                   fp_samples=fp_samples, bp_samples=bp_samples,
                                   projection_iteration=projection_iteration, skip_projections=skip_projections,
                                   voxel_type=voxel_type, init_value=init_value, x_axis_tilt=x_axis_tilt,
                                   mask=mask, flags=flags, geometries=geometries)
        self.desc = d
        self.voxel_type = d.voxel_type
        self.shape = (volume.resolution[2], volume.resolution[1], volume.resolution[0])
        h = C.c_void_p()
        call("etn_reconstruction_create", C.byref(d), C.c_void_p(stream) if stream else None, C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().etn_reconstruction_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self):
        call("etn_reconstruction_run", self._h)

    def run_single_iteration(self):
        call("etn_reconstruction_run_single_iteration", self._h)

    def run_iterations(self, n):
        call("etn_reconstruction_run_iterations", self._h, int(n))

    def synchronize(self):
        call("etn_reconstruction_synchronize", self._h)

    def residual_norm(self):
        out = C.c_double()
        call("etn_reconstruction_residual_norm", self._h, C.byref(out))
        return out.value

    def uses_fused_path(self):
        out = C.c_int()
        call("etn_reconstruction_uses_fused_path", self._h, C.byref(out))
        return bool(out.value)

    def time_kernels(self):
        """(kernel id, average ms, launches) of the dominant kernel over one extra
        eagerly launched iteration (0 fused, 1 generic FP, 2 generic BP)."""
        k = C.c_int()
        ms = C.c_double()
        n = C.c_uint32()
        call("etn_reconstruction_time_kernels", self._h, C.byref(k), C.byref(ms), C.byref(n))
        return k.value, ms.value, n.value

    def set_tile(self, planes, band):
        """Fused kernels' tile: volume planes per workgroup, z rows per LDS band."""
        call("etn_reconstruction_set_tile", self._h, int(planes), int(band))

    def tile(self):
        """(planes, band) in use; runs the autotuner first if it has not run."""
        p, b = C.c_int(), C.c_int()
        call("etn_reconstruction_tile", self._h, C.byref(p), C.byref(b))
        return p.value, b.value

    def set_workgroup_threads(self, threads):
        """Fused kernels' workgroup shape: 512 or 256 threads."""
        call("etn_reconstruction_set_workgroup_threads", self._h, int(threads))

    def workgroup_threads(self):
        """Threads per fused workgroup in use; runs the autotuner first if it has not run."""
        t = C.c_int()
        call("etn_reconstruction_workgroup_threads", self._h, C.byref(t))
        return t.value

    def volume_pointer(self):
        p = C.c_void_p()
        n = C.c_size_t()
        call("etn_reconstruction_volume", self._h, C.byref(p), C.byref(n))
        return p.value, n.value

    def get_reconstructed_volume(self, dtype=None):
        """Copy of the working volume as a device tensor [z][y][x]."""
        torch = _torch()
        tdt = {ETN_VOXEL_FLOAT32: torch.float32, ETN_VOXEL_FLOAT16: torch.float16,
               ETN_VOXEL_BFLOAT16: torch.bfloat16}[self.voxel_type]
        out = torch.empty(self.shape, dtype=dtype or tdt, device="cuda")
        self.synchronize()
        src, n = self.volume_pointer()
        count = int(np.prod(self.shape))
        call("etn_convert_volume", C.c_void_p(src), self.voxel_type, _ptr(out), _voxel_type_of(out), count,
             _stream(None))
        torch.cuda.synchronize()
        return out

    def set_volume(self, tensor):
        if _voxel_type_of(tensor) != self.voxel_type or tuple(tensor.shape) != self.shape:
            raise ValueError("volume must match the reconstruction's voxel type and shape")
        torch = _torch()
        torch.cuda.synchronize()
        call("etn_reconstruction_set_volume", self._h, _ptr(tensor))
        self.synchronize()


class WeightedBackProjection(BlockIterativeReconstructionOperator):
    """The WBP plugin's algorithm (etn_reconstruction_* with algorithm "wbp"):
    run() filters and back-projects every projection once with lambda / N;
    run_iterations(n) does projections 0..n-1."""

    def __init__(self, **kw):
        kw["algorithm"] = "wbp"
        super().__init__(**kw)


class BatchReconstruction:
    """Wraps etn_batch_*: independent reconstructions advanced together, each
    fused launch carrying one pass of every member.  `members` is a list of
    keyword dicts as for BlockIterativeReconstructionOperator (only its
    ETN_RECON_FLAG_GENERIC_KERNELS flag is honoured per member); `flags` applies
    to the whole batch.  No reference counterpart."""

    def __init__(self, members, flags=0, stream=None):
        descs, self._keep, self.shapes, self.voxel_types = [], [], [], []
        for m in members:
            d, keep = _make_desc(**m)
            descs.append(d)
            self._keep.append(keep)
            r = m["volume"].resolution
            self.shapes.append((r[2], r[1], r[0]))
            self.voxel_types.append(d.voxel_type)
        arr = (ReconstructionDesc * len(descs))(*descs)
        self._descs = arr
        h = C.c_void_p()
        call("etn_batch_create", arr, len(descs), int(flags), C.c_void_p(stream) if stream else None, C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().etn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.shapes)

    def run_iterations(self, n):
        call("etn_batch_run_iterations", self._h, int(n))

    def synchronize(self):
        call("etn_batch_synchronize", self._h)

    def residual_norm(self, i):
        out = C.c_double()
        call("etn_batch_residual_norm", self._h, int(i), C.byref(out))
        return out.value

    def uses_fused_path(self, i):
        out = C.c_int()
        call("etn_batch_uses_fused_path", self._h, int(i), C.byref(out))
        return bool(out.value)

    def tile(self, i):
        p, b = C.c_int(), C.c_int()
        call("etn_batch_tile", self._h, int(i), C.byref(p), C.byref(b))
        return p.value, b.value

    def workgroup_threads(self, i):
        t = C.c_int()
        call("etn_batch_workgroup_threads", self._h, int(i), C.byref(t))
        return t.value

    def time_iteration(self):
        """(total ms of the fused launches, launch count) over one more eager iteration."""
        ms, n = C.c_double(), C.c_uint32()
        call("etn_batch_time_iteration", self._h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def get_reconstructed_volume(self, i, dtype=None):
        """Copy of member i's volume as a device tensor [z][y][x]."""
        torch = _torch()
        vt = self.voxel_types[i]
        tdt = {ETN_VOXEL_FLOAT32: torch.float32, ETN_VOXEL_FLOAT16: torch.float16,
               ETN_VOXEL_BFLOAT16: torch.bfloat16}[vt]
        out = torch.empty(self.shapes[i], dtype=dtype or tdt, device="cuda")
        self.synchronize()
        p, n = C.c_void_p(), C.c_size_t()
        call("etn_batch_volume", self._h, int(i), C.byref(p), C.byref(n))
        count = int(np.prod(self.shapes[i]))
        call("etn_convert_volume", p, vt, _ptr(out), _voxel_type_of(out), count, _stream(None))
        torch.cuda.synchronize()
        return out
