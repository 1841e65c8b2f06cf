"""MI355X (gfx950) implementation of Ettention's block-iterative SART/SIRT hot
path: HIP kernels and a C++ host layer in libetn.so behind the C ABI of
include/etn/etn.h, with a thin Python front end (ettention_amd.api)."""
from ._lib import (ETN_ALL_SUBVOLUMES, ETN_RECON_FLAG_EXACT_ARITHMETIC, ETN_RECON_FLAG_GENERIC_KERNELS,
                   ETN_RECON_FLAG_NO_AUTOTUNE, ETN_RECON_FLAG_STOCHASTIC_ROUNDING,
                   ETN_RECON_FLAG_16BIT_PROJECTIONS, ETN_RECON_FLAG_LDS_BACK_PROJECTION,
                   ETN_RECON_FLAG_NO_GRAPH, ETN_VOXEL_BFLOAT16, ETN_VOXEL_FLOAT16, ETN_VOXEL_FLOAT32, EtnError, InvalidArgument,
                   ScannerGeometry, Unsupported, VolumeGeometry, lib)

__all__ = ["ETN_ALL_SUBVOLUMES", "ETN_RECON_FLAG_EXACT_ARITHMETIC", "ETN_RECON_FLAG_GENERIC_KERNELS", "ETN_RECON_FLAG_NO_AUTOTUNE", "ETN_RECON_FLAG_NO_GRAPH",
           "ETN_RECON_FLAG_STOCHASTIC_ROUNDING", "ETN_RECON_FLAG_16BIT_PROJECTIONS",
           "ETN_RECON_FLAG_LDS_BACK_PROJECTION",
           "ETN_VOXEL_BFLOAT16", "ETN_VOXEL_FLOAT16", "ETN_VOXEL_FLOAT32", "EtnError", "InvalidArgument",
           "ScannerGeometry", "Unsupported", "VolumeGeometry", "lib"]
