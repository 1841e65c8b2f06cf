"""ctypes binding of libetn.so (include/etn/etn.h).

The library must have been built (python -m ettention_amd.build); importing
this module without it raises ImportError -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libetn.so")

ETN_OK = 0
ETN_ERROR_INVALID_ARGUMENT = 1
ETN_ERROR_UNSUPPORTED = 2
ETN_ERROR_DEVICE = 3
ETN_ERROR_OUT_OF_MEMORY = 4
ETN_ERROR_INTERNAL = 5

ETN_VOXEL_FLOAT32 = 0
ETN_VOXEL_FLOAT16 = 1
ETN_VOXEL_BFLOAT16 = 2

ETN_ALL_SUBVOLUMES = 0xFFFFFFFF
ETN_RECON_FLAG_GENERIC_KERNELS = 1
ETN_RECON_FLAG_NO_GRAPH = 2
ETN_RECON_FLAG_NO_AUTOTUNE = 4
ETN_RECON_FLAG_EXACT_ARITHMETIC = 8
ETN_RECON_FLAG_STOCHASTIC_ROUNDING = 16
ETN_RECON_FLAG_16BIT_PROJECTIONS = 32
ETN_RECON_FLAG_LDS_BACK_PROJECTION = 64


class EtnError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("etn status %d: %s" % (status, message))
        self.status = status


class InvalidArgument(EtnError, ValueError):
    pass


class Unsupported(EtnError, NotImplementedError):
    pass


class ScannerGeometry(C.Structure):
    """etn_scanner_geometry (ParallelScannerGeometry state)."""
    _fields_ = [("source", C.c_float * 3), ("detector", C.c_float * 3),
                ("horizontal_pitch", C.c_float * 3), ("vertical_pitch", C.c_float * 3),
                ("resolution", C.c_uint32 * 2)]

    @staticmethod
    def make(source, detector, hpitch, vpitch, res):
        g = ScannerGeometry()
        g.source[:] = [float(x) for x in source]
        g.detector[:] = [float(x) for x in detector]
        g.horizontal_pitch[:] = [float(x) for x in hpitch]
        g.vertical_pitch[:] = [float(x) for x in vpitch]
        g.resolution[:] = [int(x) for x in res]
        return g

    @staticmethod
    def mrc_base(nx, ny):
        """MRCStack base geometry (MRCStack.cpp:250-262)."""
        return ScannerGeometry.make((-0.5 * nx, -0.5 * ny, float(nx)), (-0.5 * nx, -0.5 * ny, -float(nx)),
                                    (1, 0, 0), (0, 1, 0), (nx, ny))


class VolumeGeometry(C.Structure):
    """etn_volume_geometry (VolumeProperties)."""
    _fields_ = [("resolution", C.c_uint32 * 3), ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
                ("subvolume_count", C.c_uint32)]

    @staticmethod
    def make(res, bmin=None, bmax=None, subvolumes=1):
        v = VolumeGeometry()
        v.resolution[:] = [int(x) for x in res]
        if bmin is None:
            bmin = [-0.5 * r for r in res]
            bmax = [0.5 * r for r in res]
        v.bbox_min[:] = [float(x) for x in bmin]
        v.bbox_max[:] = [float(x) for x in bmax]
        v.subvolume_count = int(subvolumes)
        return v


class ReconstructionDesc(C.Structure):
    """etn_reconstruction_desc."""
    _fields_ = [("algorithm", C.c_char_p), ("block_size", C.c_uint32), ("iterations", C.c_uint32),
                ("lambda_", C.c_float), ("use_long_object_compensation", C.c_int),
                ("fp_samples", C.c_uint32), ("bp_samples", C.c_uint32),
                ("projection_iteration", C.c_char_p), ("skip_projections", C.POINTER(C.c_uint32)),
                ("skip_count", C.c_uint32), ("voxel_type", C.c_int), ("volume", VolumeGeometry),
                ("init_value", C.c_float), ("base_geometry", ScannerGeometry),
                ("projection_count", C.c_uint32), ("tilt_angles", C.POINTER(C.c_float)),
                ("x_axis_tilt", C.c_float), ("projections", C.c_void_p), ("mask", C.c_void_p),
                ("flags", C.c_uint32), ("geometries", C.POINTER(ScannerGeometry))]


# (name, restype, argtypes) of every entry point declared in include/etn/etn.h
_vp, _u32, _f32, _i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
_SG, _VG = C.POINTER(ScannerGeometry), C.POINTER(VolumeGeometry)
SIGNATURES = [
    ("etn_last_error", C.c_char_p, []),
    ("etn_version", C.c_char_p, []),
    ("etn_rotate_scanner_geometry", _i32, [_SG, _f32, _f32, _SG]),
    ("etn_ray_direction", _i32, [_SG, _vp]),
    ("etn_projection_matrix", _i32, [_SG, _vp]),
    ("etn_traversal_length_through_bbox", _i32, [_SG, _vp, _vp, _vp]),
    ("etn_bp_sample_pattern", _i32, [_SG, _vp, _u32, _vp]),
    ("etn_subvolume", _i32, [_VG, _u32, _vp, _vp, _vp, _vp]),
    ("etn_projection_order", _i32, [C.c_char_p, _u32, _vp, _u32, _vp, _vp]),
    ("etn_forward_project", _i32, [_VG, _i32, _vp, _vp, _u32, _SG, _u32, _vp, _vp, _vp]),
    ("etn_back_project", _i32, [_VG, _i32, _vp, _vp, _u32, _SG, _u32, _vp, _vp, _f32, _u32, _vp]),
    ("etn_compare", _i32, [_u32, _u32, _vp, _vp, _vp, _vp]),
    ("etn_long_object_compensate", _i32, [_u32, _u32, _vp, _vp, _f32, _vp]),
    ("etn_compute_ray_length", _i32, [_VG, _SG, _vp, _vp]),
    ("etn_ramp_filter", _i32, [_u32, _u32, _vp, _vp, _vp]),
    ("etn_sum_squared_difference_f64", _i32, [_vp, _vp, C.c_size_t, _vp, _vp]),
    ("etn_sum_squared_difference_f64_noalloc", _i32, [_vp, _vp, C.c_size_t, _vp, _vp]),
    ("etn_convert_volume", _i32, [_vp, _i32, _vp, _i32, C.c_size_t, _vp]),
    ("etn_reconstruction_create", _i32, [C.POINTER(ReconstructionDesc), _vp, C.POINTER(_vp)]),
    ("etn_reconstruction_destroy", _i32, [_vp]),
    ("etn_reconstruction_run", _i32, [_vp]),
    ("etn_reconstruction_run_single_iteration", _i32, [_vp]),
    ("etn_reconstruction_run_iterations", _i32, [_vp, _u32]),
    ("etn_reconstruction_volume", _i32, [_vp, C.POINTER(_vp), C.POINTER(C.c_size_t)]),
    ("etn_reconstruction_set_volume", _i32, [_vp, _vp]),
    ("etn_reconstruction_residual_norm", _i32, [_vp, C.POINTER(C.c_double)]),
    ("etn_reconstruction_synchronize", _i32, [_vp]),
    ("etn_reconstruction_uses_fused_path", _i32, [_vp, C.POINTER(C.c_int)]),
    ("etn_reconstruction_time_kernels", _i32, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                               C.POINTER(C.c_uint32)]),
    ("etn_reconstruction_set_tile", _i32, [_vp, C.c_int, C.c_int]),
    ("etn_reconstruction_tile", _i32, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("etn_reconstruction_set_workgroup_threads", _i32, [_vp, C.c_int]),
    ("etn_reconstruction_workgroup_threads", _i32, [_vp, C.POINTER(C.c_int)]),
    ("etn_batch_create", _i32, [C.POINTER(ReconstructionDesc), _u32, _u32, _vp, C.POINTER(_vp)]),
    ("etn_batch_destroy", _i32, [_vp]),
    ("etn_batch_run_iterations", _i32, [_vp, _u32]),
    ("etn_batch_synchronize", _i32, [_vp]),
    ("etn_batch_volume", _i32, [_vp, _u32, C.POINTER(_vp), C.POINTER(C.c_size_t)]),
    ("etn_batch_residual_norm", _i32, [_vp, _u32, C.POINTER(C.c_double)]),
    ("etn_batch_uses_fused_path", _i32, [_vp, _u32, C.POINTER(C.c_int)]),
    ("etn_batch_tile", _i32, [_vp, _u32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("etn_batch_workgroup_threads", _i32, [_vp, _u32, C.POINTER(C.c_int)]),
    ("etn_batch_time_iteration", _i32, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
    ("etn_unet_create", _i32, [_u32, _u32, _u32, _vp, C.POINTER(_vp)]),
    ("etn_unet_destroy", _i32, [_vp]),
    ("etn_unet_parameter_count", _i32, [C.POINTER(C.c_size_t)]),
    ("etn_unet_set_parameters", _i32, [_vp, _vp, C.c_size_t]),
    ("etn_unet_get_parameters", _i32, [_vp, _vp, C.c_size_t]),
    ("etn_unet_predict", _i32, [_vp, _vp, _vp, _f32, _f32, _vp]),
    ("etn_unet_gradient", _i32, [_vp, _vp, _vp, _vp, _f32, _f32, _vp, C.POINTER(C.c_double)]),
    ("etn_unet_train_step", _i32, [_vp, _vp, _vp, _vp, _f32, _f32, _f32, C.POINTER(C.c_double)]),
    ("etn_unet_time_forward", _i32, [_vp, _vp, _f32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("etn_projection_stack_open", _i32, [C.c_char_p, C.c_char_p, _i32, C.POINTER(_vp)]),
    ("etn_projection_stack_destroy", _i32, [_vp]),
    ("etn_projection_stack_info", _i32, [_vp, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32), _SG]),
    ("etn_projection_stack_read", _i32, [_vp, _vp, _vp]),
    ("etn_image_stack_directory_write", _i32, [C.c_char_p, _vp, _u32, _u32, _u32, _vp, _SG, _vp]),
    ("etn_tiff_read", _i32, [C.c_char_p, _u32, C.POINTER(_u32), C.POINTER(_u32), _vp]),
    ("etn_tiff_page_count", _i32, [C.c_char_p, C.POINTER(_u32)]),
    ("etn_tiff_write", _i32, [C.c_char_p, _vp, _u32, _u32]),
    ("etn_mrc_write_volume", _i32, [C.c_char_p, _vp, _vp, C.c_char_p, C.c_char_p, _i32]),
    ("etn_mrc_read_volume", _i32, [C.c_char_p, C.c_char_p, _vp, _vp]),
    ("etn_config_resolve", _i32, [_i32, C.POINTER(C.c_char_p), C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("etn_main", _i32, [_i32, C.POINTER(C.c_char_p)]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libetn.so is not built (run python -m ettention_amd.build); "
                              "there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status == ETN_OK:
        return
    msg = lib().etn_last_error().decode(errors="replace")
    if status == ETN_ERROR_INVALID_ARGUMENT:
        raise InvalidArgument(status, msg)
    if status == ETN_ERROR_UNSUPPORTED:
        raise Unsupported(status, msg)
    raise EtnError(status, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))
