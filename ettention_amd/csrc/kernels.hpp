// Parameter blocks passed by value to the projector kernels, and the host
// launchers defined in projectors.hip.  Shared by host and device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "etn/etn.h"

namespace etn {
namespace dev {

constexpr int kMaxBpSamples = 64;

// VolumeGeometry of one subvolume (ocl_src/VolumeGeometry.hl:4-15) plus the
// strides of the complete volume the subvolume lives in.
struct SlabParams {
    float bmin[3], bmax[3];     // subvolume bounding box
    float vs[3];                // voxel size
    int res[3];                 // subvolume resolution
    uint32_t nx, ny;            // complete-volume strides
    uint64_t first_voxel;       // linear index of the subvolume's first voxel
};

// ProjectionGeometry fields the FP reads (ocl_src/ProjectionGeometry.hl:4-15).
struct RayParams {
    float source[3], h[3], v[3], ray[3];
    uint32_t nu, nv, samples;
};

// Inputs of ParallelBeamsBackProjection.cl:correct besides the volume.
struct BpParams {
    float m[16];                // projection matrix, column-major
    float lambda;               // lambda / block size
    uint32_t nu, nv, nsamples;
    float4 samples[kMaxBpSamples];   // xyz offset, w = in-voxel chord length
};

void launch_forward_project(const SlabParams& s, const RayParams& r, etn_voxel_type vt, const void* volume,
                            const uint8_t* mask, float* projection, float* traversal_length, hipStream_t stream);
void launch_back_project(const SlabParams& s, const BpParams& b, etn_voxel_type vt, void* volume,
                         const uint8_t* mask, const float* residual, const float* ray_length, hipStream_t stream);
void launch_compare(uint32_t n, const float* real, const float* virt, float* residual, hipStream_t stream);
// real in its storage type (measured projections stored as f16 or bf16)
void launch_compare(uint32_t n, const void* real, etn_voxel_type real_type, const float* virt, float* residual,
                    hipStream_t stream);
void launch_long_object_compensate(uint32_t n, float* projection, float* length, float c, hipStream_t stream);
void launch_ray_length(const float bmin[3], const float bmax[3], const RayParams& r, float* out, hipStream_t stream);
// `count` images of nu x nv rays in one launch: image i with the geometry
// params[i] (device memory, every nu and nv the same) at out + i * stride
void launch_ray_length_batch(const float bmin[3], const float bmax[3], const RayParams* params, uint32_t count,
                             uint32_t nu, uint32_t nv, float* out, size_t stride, hipStream_t stream);
// partials must hold ssd_partial_count() doubles
constexpr uint32_t kSsdBlocks = 512;
void launch_sum_squared_difference(const float* a, const float* b, size_t count, double* partials, double* out,
                                   hipStream_t stream);
// a in its storage type
void launch_sum_squared_difference(const void* a, etn_voxel_type a_type, const float* b, size_t count,
                                   double* partials, double* out, hipStream_t stream);
// Device address of `device`'s static partial-sum array (kSsdBlocks doubles),
// looked up once per device with that device current.
double* ssd_static_scratch(int device);
void launch_convert(const void* src, etn_voxel_type st, void* dst, etn_voxel_type dt, size_t count, hipStream_t stream);
void launch_fill(void* dst, etn_voxel_type t, float value, size_t count, hipStream_t stream);

// Ramp filtering of one projection as the reference's Fourier-space
// convolution computes it (ConvolutionFourierSpaceImplementation.cpp:50-83):
// the image mirror-padded to ux x uy (powers of two), convolved circularly
// with `taps` filter values centred on tap taps / 2, the filter one padded row
// above the origin, cropped back to nu x nv.
struct RampFilterParams {
    uint32_t nu, nv;            // image
    uint32_t ux, uy;            // padded size
    int32_t pad_x, pad_y;       // first image column / row inside the padded image
    int32_t end_x, end_y;       // one past the last
    int32_t crop_x, crop_y;     // first padded column / row of the cropped result
    uint32_t taps;
};
constexpr uint32_t kMaxRampRow = 8192;    // padded row length held in LDS
constexpr uint32_t kMaxRampTaps = 512;
void launch_ramp_filter(const RampFilterParams& p, const float* taps, const float* src, float* dst,
                        hipStream_t stream);
// `count` images in one launch: image i from src + i * src_stride to
// dst + i * dst_stride (floats), each filtered as launch_ramp_filter does.
void launch_ramp_filter_batch(const RampFilterParams& p, const float* taps, const float* src, size_t src_stride,
                              float* dst, size_t dst_stride, uint32_t count, hipStream_t stream);
void launch_zero(float* dst, size_t count, hipStream_t stream);

}  // namespace dev
}  // namespace etn
