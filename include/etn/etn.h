/*
 * etn.h -- C ABI of libetn.so, the MI355X (gfx950) implementation of
 * Ettention's block-iterative SART/SIRT hot path.
 *
 * Conventions
 *   - Plain C types only.  Device pointers are HIP device addresses; `stream`
 *     is a hipStream_t passed as void* (NULL = the null stream).
 *   - Every function returns an etn_status; on failure etn_last_error()
 *     returns a message (thread-local, valid until the next call on the thread).
 *   - Kernel entry points are stream-ordered and asynchronous; they allocate
 *     nothing on the device except where stated.
 *   - Volumes are stored as in the reference: linear index z*ny*nx + y*nx + x
 *     over the complete volume; a z-slab ("subvolume") is a contiguous run of
 *     z slices (VolumeProperties.cpp:74-98).
 *
 * Reference paths are relative to the reference checkout root.
 */
#ifndef ETN_ETN_H
#define ETN_ETN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum etn_status {
    ETN_OK = 0,
    ETN_ERROR_INVALID_ARGUMENT = 1, /* precondition the reference throws Exception for */
    ETN_ERROR_UNSUPPORTED = 2,      /* no kernel variant for this geometry / basis */
    ETN_ERROR_DEVICE = 3,           /* HIP runtime error (reference: GPUException / CL_ASSERT) */
    ETN_ERROR_OUT_OF_MEMORY = 4,
    ETN_ERROR_INTERNAL = 5
} etn_status;

/* Voxel storage (optimization.internalVoxelRepresentationType; Voxel.cpp:45-72).
 * Arithmetic is FP32 for every storage type. */
typedef enum etn_voxel_type {
    ETN_VOXEL_FLOAT32 = 0,
    ETN_VOXEL_FLOAT16 = 1,          /* IEEE half, round-to-nearest-even on store (vstore_half) */
    ETN_VOXEL_BFLOAT16 = 2          /* extension: bfloat16, round-to-nearest-even on store */
} etn_voxel_type;

/* State of a ParallelScannerGeometry (ScannerGeometry.h; ScannerGeometry.cpp:95-103). */
typedef struct etn_scanner_geometry {
    float source[3];
    float detector[3];
    float horizontal_pitch[3];
    float vertical_pitch[3];
    uint32_t resolution[2];         /* detector (nu, nv) */
} etn_scanner_geometry;

/* VolumeProperties: resolution, complete bounding box, z-slab count
 * (VolumeProperties.cpp:21-98; VolumeParameterSet.cpp:68-71). */
typedef struct etn_volume_geometry {
    uint32_t resolution[3];
    float bbox_min[3];
    float bbox_max[3];
    uint32_t subvolume_count;       /* >= 1 */
} etn_volume_geometry;

#define ETN_ALL_SUBVOLUMES 0xffffffffu   /* BackProjectionOperator::ALL_SUB_VOLUMES */

const char* etn_last_error(void);
const char* etn_version(void);

/* ---------------------------------------------------------------------------
 * Host-side geometry (no device work)
 * ------------------------------------------------------------------------- */

/* SingleAxisTiltRotator::createRotatedScannerGeometry
 * (ettention/libReconstruction/src/model/geometry/SingleAxisTiltRotator.cpp:42-78). */
etn_status etn_rotate_scanner_geometry(const etn_scanner_geometry* base, float tilt_deg,
                                       float x_axis_tilt_deg, etn_scanner_geometry* out);

/* ParallelScannerGeometry::getRayDirection / getProjectionMatrix (column-major,
 * values[4*col+row]) / getTraversalLengthThroughBoundingBox
 * (ettention/libReconstruction/src/model/geometry/ParallelScannerGeometry.cpp:44-69). */
etn_status etn_ray_direction(const etn_scanner_geometry* g, float out[3]);
etn_status etn_projection_matrix(const etn_scanner_geometry* g, float out[16]);
etn_status etn_traversal_length_through_bbox(const etn_scanner_geometry* g, const float bbox_min[3],
                                             const float bbox_max[3], float* out);

/* ParallelBeamsBackProjectionKernel::generateSamplingPatternForPerVoxelSampling
 * (ParallelBeamsBackProjectionKernel.cpp:69-103); out4 holds samples x (dx, dy, dz, l).
 * Fails for samples that are zero or not a perfect square (:55-63). */
etn_status etn_bp_sample_pattern(const etn_scanner_geometry* g, const float voxel_size[3],
                                 uint32_t samples, float* out4);

/* VolumeProperties::setSubVolumeCount / getSubVolume* (VolumeProperties.cpp:74-98). */
etn_status etn_subvolume(const etn_volume_geometry* v, uint32_t index, uint32_t resolution_out[3],
                         uint32_t* first_slice, float bbox_min_out[3], float bbox_max_out[3]);

/* ProjectionIteratorFactory::instantiate + ProjectionSet::setProjectionCount/reorder:
 * scheme "identity" | "random" | "maxangle" | "buckets"
 * (ettention/libReconstruction/src/model/projectionset/factory/ProjectionIteratorFactory.cpp:11-23).
 * `out` must hold `count` entries. */
etn_status etn_projection_order(const char* scheme, uint32_t count, const uint32_t* skip,
                                uint32_t skip_count, uint32_t* out, uint32_t* out_count);

/* ---------------------------------------------------------------------------
 * Kernels (device pointers, stream-ordered)
 * ------------------------------------------------------------------------- */

/* ParallelBeamsForwardProjectionKernel::run for one subvolume: DDA line
 * integral of samples^2 sub-rays per pixel, ADDED to projection and
 * traversal_length (ForwardProjectionOperator::Execute,
 * ettention/libReconstruction/src/algorithm/projections/forward/ForwardProjectionOperator.cpp:31-50;
 * ocl_src/ParallelBeamsForwardProjection.cl:26-58).  `volume` points at the
 * complete volume; `mask` (uint8, 0..255, complete volume) may be NULL.
 * samples 0 is treated as 1 (ForwardProjectionKernel.cpp:65-68). */
etn_status etn_forward_project(const etn_volume_geometry* vol, etn_voxel_type voxel_type,
                               const void* volume, const uint8_t* mask, uint32_t subvolume,
                               const etn_scanner_geometry* g, uint32_t samples,
                               float* projection, float* traversal_length, void* stream);

/* BackProjectionOperator::setInput(residual, rayLength, blockSize) + execute(subvolume)
 * (ettention/libReconstruction/src/algorithm/projections/back/BackProjectionOperator.cpp:43-98;
 * ocl_src/ParallelBeamsBackProjection.cl:9-53): lambda is divided by block_size;
 * subvolume may be ETN_ALL_SUBVOLUMES. */
etn_status etn_back_project(const etn_volume_geometry* vol, etn_voxel_type voxel_type, void* volume,
                            const uint8_t* mask, uint32_t subvolume, const etn_scanner_geometry* g,
                            uint32_t samples, const float* residual, const float* ray_length,
                            float lambda, uint32_t block_size, void* stream);

/* CompareKernel: residual = real - virtual
 * (ettention/libReconstruction/src/algorithm/imagemanipulation/CompareKernel.cpp:9-66; ocl_src/Compare.cl). */
etn_status etn_compare(uint32_t nu, uint32_t nv, const float* real, const float* virtual_projection,
                       float* residual, void* stream);

/* VirtualProjectionKernel (long-object compensation): p = L > 1e-4 ? p*c/L : 0; L = c
 * (ettention/libReconstruction/src/algorithm/longobjectcompensation/LongObjectCompensationKernel.cpp:9-59;
 * ocl_src/VirtualProjection.cl). */
etn_status etn_long_object_compensate(uint32_t nu, uint32_t nv, float* projection,
                                      float* traversal_length, float constant_length, void* stream);

/* ComputeRayLengthKernel: chord of each pixel-centre ray through the complete
 * bounding box (ComputeRayLengthKernel.cpp:9-46; ocl_src/ComputeRayLength.cl). */
etn_status etn_compute_ray_length(const etn_volume_geometry* vol, const etn_scanner_geometry* g,
                                  float* ray_length, void* stream);
/* ConvolutionOperator with the 128-tap RamLak kernel as WBP sets it up
 * (WBP.cpp:29-33; GenerateRamLakKernel.cl; ConvolutionFourierSpaceImplementation.cpp):
 * the image mirror-padded to powers of two, convolved circularly along rows
 * with the filter placed one padded row above the origin, cropped back.
 * Device images of nu x nv floats; synchronous on `stream`. */
etn_status etn_ramp_filter(uint32_t nu, uint32_t nv, const float* projection, float* filtered, void* stream);

/* Extension (no reference counterpart; the reference's only metric is the racy
 * BufferRMS.cl): out_device[0] += sum_i (a_i - b_i)^2 accumulated in FP64, with b
 * = NULL meaning zero.  Deterministic for a given count. */
etn_status etn_sum_squared_difference_f64(const float* a, const float* b, size_t count,
                                          double* out_device, void* stream);

/* Allocation-free variant of etn_sum_squared_difference_f64, with the same
 * kernels and the same result bit for bit.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its FP64
 * partial sums live in a static device array of ETN_SSD_NOALLOC_SCRATCH_BYTES
 * per GPU (its address is looked up on the first call on each GPU), so it
 * calls no hipMalloc/hipFree.  The partial count is fixed, so the scratch
 * bounds nothing but count <= ETN_SSD_NOALLOC_MAX_COUNT (2^40 elements, 4 TiB
 * of floats; larger counts are refused).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (different streams, host
 * threads) race, so serialise them. */
#define ETN_SSD_NOALLOC_SCRATCH_BYTES 4096u
#define ETN_SSD_NOALLOC_MAX_COUNT ((size_t)1 << 40)
etn_status etn_sum_squared_difference_f64_noalloc(const float* a, const float* b, size_t count,
                                                  double* out_device, void* stream);

/* Storage conversion helpers for reduced-precision volumes (device to device). */
etn_status etn_convert_volume(const void* src, etn_voxel_type src_type, void* dst,
                              etn_voxel_type dst_type, size_t count, void* stream);

/* ---------------------------------------------------------------------------
 * Block-iterative reconstruction (BlockIterativeReconstructionOperator)
 * ------------------------------------------------------------------------- */

typedef struct etn_reconstruction etn_reconstruction;

enum {
    ETN_RECON_FLAG_GENERIC_KERNELS = 1u << 0,   /* disable the fused plane kernels */
    ETN_RECON_FLAG_NO_GRAPH = 1u << 1,          /* launch eagerly instead of replaying a hipGraph */
    ETN_RECON_FLAG_NO_AUTOTUNE = 1u << 2,       /* keep the default fused tile instead of timing candidates */
    ETN_RECON_FLAG_EXACT_ARITHMETIC = 1u << 3   /* fused kernels repeat the generic kernels' float operations in
                                                   their order (bit-identical volumes); by default they combine
                                                   the bilinear weights per detector column, use the hardware
                                                   reciprocal for chord / L and fused multiply-adds */,
    ETN_RECON_FLAG_STOCHASTIC_ROUNDING = 1u << 4 /* 16-bit storage: round each back-projected voxel to storage
                                                   stochastically (up with probability equal to the dropped
                                                   fraction; random bits hashed from the voxel index, the
                                                   projection and the value, so runs are reproducible) instead
                                                   of to nearest even.  Extension (the reference rounds to
                                                   nearest); needs the fast fused kernels: f16 or bf16 storage,
                                                   no mask, no oversampling, nx a multiple of 4, rays that
                                                   cannot reach the 2048-step limit (nx + nz <= 2049 always
                                                   qualifies);
                                                   creation fails otherwise */,
    ETN_RECON_FLAG_16BIT_PROJECTIONS = 1u << 5,  /* the measured projections are stored in the volume's 16-bit
                                                   type (voxel_type f16 or bf16): `projections` addresses
                                                   projection_count * nu * nv such values.  Arithmetic stays
                                                   FP32.  Extension (the reference keeps projections in f32);
                                                   not for "wbp" */
    ETN_RECON_FLAG_LDS_BACK_PROJECTION = 1u << 6 /* fast fused kernels: apply the back projections to the band in
                                                   LDS, one projection after another, instead of in registers as
                                                   the band arrives from HBM (the default where it applies: nx a
                                                   multiple of 4, no oversampling).  Same volumes bit for bit;
                                                   for comparison and measurement */
};

/* Parameters of BlockIterativeReconstructionOperator, named after the
 * reference keys (AlgorithmParameterSet.cpp:8-53, OptimizationParameterSet.cpp:7-22,
 * ForwardProjectionParameterSet, BackprojectionParameterSet, VolumeParameterSet,
 * InputParameterSet, HardwareParameterSet). */
typedef struct etn_reconstruction_desc {
    const char* algorithm;              /* "sart" | "sirt" | "blockiterative" | "wbp" (case-insensitive) */
    uint32_t block_size;                /* algorithm.blockSize (blockiterative only) */
    uint32_t iterations;                /* algorithm.numberOfIterations (run) */
    float lambda;                       /* algorithm.lambda */
    int use_long_object_compensation;   /* algorithm.useLongObjectCompensation */
    uint32_t fp_samples;                /* forwardProjection.samples */
    uint32_t bp_samples;                /* backProjection.samples (perfect square) */
    const char* projection_iteration;   /* optimization.projectionIteration; NULL = "random" */
    const uint32_t* skip_projections;   /* input.skipProjections */
    uint32_t skip_count;
    etn_voxel_type voxel_type;          /* optimization.internalVoxelRepresentationType */
    etn_volume_geometry volume;         /* volume.* + hardware.subVolumeCount (0 = 1 on this GPU) */
    float init_value;                   /* volume.initValue */
    etn_scanner_geometry base_geometry; /* untilted geometry of the projection stack */
    uint32_t projection_count;
    const float* tilt_angles;           /* host, degrees, projection_count entries */
    float x_axis_tilt;                  /* input.xAxisTilt */
    const void* projections;            /* device, projection_count * nu * nv floats in stack order
                                           (16-bit values with ETN_RECON_FLAG_16BIT_PROJECTIONS) */
    const uint8_t* mask;                /* device prior-knowledge mask or NULL */
    uint32_t flags;                     /* ETN_RECON_FLAG_* */
    /* host, projection_count entries, or NULL: the geometry of each projection
     * as ImageStackDatasource::createScannerGeometry returns it
     * (ImageStackDatasource.h:26; MRCStack.cpp:281-293,
     * ImageStackDirectoryDatasource.cpp:268-276).  When given, it replaces
     * rotating base_geometry by tilt_angles and x_axis_tilt, and tilt_angles may
     * be NULL; base_geometry.resolution still gives the detector size. */
    const etn_scanner_geometry* geometries;
} etn_reconstruction_desc;

/* The block-iterative ids select BlockIterativeReconstructionOperator
 * (Framework.cpp:142-152).  "wbp" selects weighted back projection
 * (plugins/WBPPlugin/src/reconstruction/WBP.cpp:13-84), which reads
 * lambda, bp_samples, volume, init_value, voxel_type, the stack, its tilt
 * angles and x_axis_tilt; run() filters and back-projects every projection
 * once, run_iterations(n) projections 0..n-1 (WBP::runNiterations).  The
 * block-iterative extensions (residual norm, kernel timing, tiles) return
 * ETN_ERROR_UNSUPPORTED for it. */
etn_status etn_reconstruction_create(const etn_reconstruction_desc* desc, void* stream,
                                     etn_reconstruction** out);
etn_status etn_reconstruction_destroy(etn_reconstruction* r);
/* ReconstructionAlgorithm::run / runSingleIteration (ReconstructionAlgorithm.h:9-22);
 * asynchronous on the reconstruction's stream. */
etn_status etn_reconstruction_run(etn_reconstruction* r);
etn_status etn_reconstruction_run_single_iteration(etn_reconstruction* r);
etn_status etn_reconstruction_run_iterations(etn_reconstruction* r, uint32_t iterations);
/* getReconstructedVolume: library-owned device buffer of the working volume. */
etn_status etn_reconstruction_volume(etn_reconstruction* r, void** device_volume, size_t* bytes);
/* Replace the working volume (volume.initFile): device source of the same type and size. */
etn_status etn_reconstruction_set_volume(etn_reconstruction* r, const void* device_volume);
/* ||A x - b||_2 over every used projection, FP64 accumulation; synchronises.
 * On the fused path the forward projections run as forward-only fused passes
 * (bit-identical to the generic kernels' norm with exact arithmetic). */
etn_status etn_reconstruction_residual_norm(etn_reconstruction* r, double* out);
etn_status etn_reconstruction_synchronize(etn_reconstruction* r);
/* 1 when the fused plane kernels run this reconstruction, 0 for the generic kernels. */
etn_status etn_reconstruction_uses_fused_path(etn_reconstruction* r, int* out);
/* Extension (profiling; the reference's CL-event PerformanceReport,
 * CLAbstractionLayer.cpp:295-329): runs ONE more iteration eagerly with HIP
 * events around every launch on the reconstruction's stream and reports the
 * average duration of the kernel that dominates it: *kernel = 0 fused plane
 * kernel (average over the launches carrying a back and a forward projection;
 * *launches = every launch of the iteration), 1 generic forward projection,
 * 2 generic back projection (*launches = that kernel's launches).  Advances
 * the reconstruction. */
etn_status etn_reconstruction_time_kernels(etn_reconstruction* r, int* kernel, double* average_ms,
                                           uint32_t* launches);
/* Extension (no reference counterpart): the fused kernels' tile, i.e. volume
 * planes per workgroup and z rows per LDS band.  The first run times the
 * candidate tiles on the reconstruction's own data unless
 * ETN_RECON_FLAG_NO_AUTOTUNE is set; setting a tile skips that.  Tiles never
 * change results.  ETN_ERROR_UNSUPPORTED on the generic path,
 * ETN_ERROR_INVALID_ARGUMENT for a tile that does not fit (planes must be
 * 1, 2, 4 or 8 and divide ny; band and detector rows must fit the LDS array
 * of the workgroup shape: 64 KiB at 512 threads, 40 KiB at 256). */
etn_status etn_reconstruction_set_tile(etn_reconstruction* r, int planes, int band);
etn_status etn_reconstruction_tile(etn_reconstruction* r, int* planes, int* band);
/* Extension (no reference counterpart): threads per workgroup of the fused
 * kernels, 512 (two workgroups per CU) or 256 (four per CU, half the band).
 * The autotuner chooses it with the tile; setting it skips the autotuner, and
 * a tile that no longer fits reverts to the default tile.  Never changes
 * results.  ETN_ERROR_UNSUPPORTED on the generic path,
 * ETN_ERROR_INVALID_ARGUMENT for another size or one the volume does not fit. */
etn_status etn_reconstruction_set_workgroup_threads(etn_reconstruction* r, int threads);
etn_status etn_reconstruction_workgroup_threads(etn_reconstruction* r, int* threads);

/* ---------------------------------------------------------------------------
 * Batched reconstructions (extension, no reference counterpart): many
 * independent reconstructions of different sizes advanced together.  Each
 * fused launch carries one pass of every member that still has one, so small
 * volumes share the GPU instead of running one after another; members with no
 * fused form run their own kernels in the same captured iteration.
 * ------------------------------------------------------------------------- */

typedef struct etn_batch etn_batch;

/* members[i] as for etn_reconstruction_create (flags there: only
 * ETN_RECON_FLAG_GENERIC_KERNELS is honoured per member); `flags` applies to the
 * whole batch.  All members share `stream` (NULL: a library-owned stream). */
etn_status etn_batch_create(const etn_reconstruction_desc* members, uint32_t count, uint32_t flags, void* stream,
                            etn_batch** out);
etn_status etn_batch_destroy(etn_batch* b);
/* `iterations` iterations of every member, asynchronous on the batch stream. */
etn_status etn_batch_run_iterations(etn_batch* b, uint32_t iterations);
etn_status etn_batch_synchronize(etn_batch* b);
etn_status etn_batch_volume(etn_batch* b, uint32_t member, void** device_volume, size_t* bytes);
etn_status etn_batch_residual_norm(etn_batch* b, uint32_t member, double* out);
etn_status etn_batch_uses_fused_path(etn_batch* b, uint32_t member, int* out);
etn_status etn_batch_tile(etn_batch* b, uint32_t member, int* planes, int* band);
/* Threads per workgroup of the fused launches that carry `member`. */
etn_status etn_batch_workgroup_threads(etn_batch* b, uint32_t member, int* threads);
/* One more eager iteration of every member with events around every fused
 * launch: total milliseconds of those launches and their count.  Advances the batch. */
etn_status etn_batch_time_iteration(etn_batch* b, double* fused_ms, uint32_t* launches);

/* ---------------------------------------------------------------------------
 * Learned correction (extension, no reference counterpart): a 3-level 3-D
 * U-Net (16/32/64 channels, 3x3x3 convolutions, bf16 on the matrix cores with
 * FP32 accumulation) mapping an early SART iterate x and its last update d (x
 * minus the iterate before it) to a late iterate, x + delta(x, d).  Volumes
 * are device float32 [nz][ny][nx]; sizes are multiples of 4.  The network
 * sees (x scale, d update_scale) and its correction is divided by scale; d
 * may be NULL (read as zeros).  The flat parameter vector holds, per layer c1a c1b c2a
 * c2b c3a c3b c4a c4b c5a c5b, weights [cout][27][cin] then biases, then the
 * head's 16 weights and bias.
 * ------------------------------------------------------------------------- */

typedef struct etn_unet etn_unet;

etn_status etn_unet_create(uint32_t nx, uint32_t ny, uint32_t nz, void* stream, etn_unet** out);
etn_status etn_unet_destroy(etn_unet* n);
etn_status etn_unet_parameter_count(size_t* out);
etn_status etn_unet_set_parameters(etn_unet* n, const float* host, size_t count);
etn_status etn_unet_get_parameters(etn_unet* n, float* host, size_t count);
/* out = x + delta(x scale, d update_scale) / scale; asynchronous on the network's stream. */
etn_status etn_unet_predict(etn_unet* n, const float* x, const float* d, float scale, float update_scale,
                            float* out);
/* Forward + backward of L = mean((x scale + delta - target scale)^2); copies the
 * gradient (FP32, parameter order) to `grad_host` when not NULL. Synchronises. */
etn_status etn_unet_gradient(etn_unet* n, const float* x, const float* d, const float* target, float scale,
                             float update_scale, float* grad_host, double* loss);
/* etn_unet_gradient followed by one Adam step with learning rate `lr`. */
etn_status etn_unet_train_step(etn_unet* n, const float* x, const float* d, const float* target, float scale,
                               float update_scale, float lr, double* loss);
/* HIP-event time of one forward pass and its convolution flops. */
etn_status etn_unet_time_forward(etn_unet* n, const float* x, float scale, double* ms, double* flops);

/* ---------------------------------------------------------------------------
 * Input / output formats and the command-line reconstruction.  Host memory
 * only; every array is caller-owned.
 * ------------------------------------------------------------------------- */

typedef struct etn_projection_stack etn_projection_stack;

/* ImageStackDatasourceFactory::instantiate (io/ImageStackDataSourceFactory.cpp):
 * a directory holding metadata.xml (ImageStackDirectoryDatasource.cpp:181-276)
 * or an .mrc / .ali stack (MRCStack.cpp:136-263, modes 0/1/2/6, optional
 * log(dataMax / (raw - min)) when `logarithmize`).  `tilt_angles` (a .tlt file,
 * ImageStackDatasource.cpp:40-73) may be NULL or "" for an MRC stack (all
 * angles 0) and is not used for a directory. */
etn_status etn_projection_stack_open(const char* projections, const char* tilt_angles, int logarithmize,
                                     etn_projection_stack** out);
etn_status etn_projection_stack_destroy(etn_projection_stack* s);
/* Detector size, projection count and the untilted base geometry. */
etn_status etn_projection_stack_info(const etn_projection_stack* s, uint32_t* nu, uint32_t* nv, uint32_t* count,
                                     etn_scanner_geometry* base);
/* Copies the images ([count][nv][nu], stack order) and tilt angles (degrees);
 * either pointer may be NULL. */
etn_status etn_projection_stack_read(const etn_projection_stack* s, float* projections, float* tilt_angles);
/* ImageStackDirectoryDatasource::writeMetaDataFile
 * (ImageStackDirectoryDatasource.cpp:134-178) plus one float TIFF per
 * projection, projection_<i>.tif, into an existing directory. */
etn_status etn_image_stack_directory_write(const char* directory, const float* projections, uint32_t count,
                                           uint32_t nu, uint32_t nv, const float* tilt_angles,
                                           const etn_scanner_geometry* base, const uint32_t volume_resolution[3]);

/* ImageDeserializer::readImage / readImageFromStack (ImageDeserializer.cpp):
 * grey TIFF (float32, uint8 / 255, uint16 / 65535), row 0 of the file is row
 * y = 0.  With pixels == NULL only the size is returned. */
etn_status etn_tiff_read(const char* path, uint32_t page, uint32_t* width, uint32_t* height, float* pixels);
etn_status etn_tiff_page_count(const char* path, uint32_t* pages);
/* Uncompressed little-endian float32 TIFF. */
etn_status etn_tiff_write(const char* path, const float* pixels, uint32_t width, uint32_t height);

/* VolumeFileFormatWriter::write + MRCWriter (VolumeFileFormatWriter.cpp:11-144,
 * MRCWriter.cpp:43-80).  `volume` is [z][y][x] with resolution (nx, ny, nz);
 * voxel_type "unsigned8" | "signed16" | "unsigned16" | "float32"; orientation
 * "xzy_order" (the reference default) | "xyz_order" | "yxz_order" |
 * "yzx_order" | "zxy_order" | "zyx_order". */
etn_status etn_mrc_write_volume(const char* path, const float* volume, const uint32_t resolution[3],
                                const char* voxel_type, const char* orientation, int invert);
/* VolumeDeserializer::load + applyInitialVolumeToWorkingVolume
 * (BlockIterativeConfiguration.cpp:156-196): the file's axes mapped by
 * `orientation`.  With volume == NULL only the resolution is returned. */
etn_status etn_mrc_read_volume(const char* path, const char* orientation, uint32_t resolution[3], float* volume);

/* The parameter sets a command line (and the --config XML file it names)
 * resolves to, as one JSON object; *length receives the size without the
 * terminating zero (CascadingParameterSource + AlgebraicParameterSet). */
etn_status etn_config_resolve(int argc, const char* const* argv, char* json, size_t capacity, size_t* length);
/* main() of the command-line tool (frontend/src/main.cpp:71-140): parses
 * argv, reconstructs on the GPU and writes the output volume; returns the
 * process exit code (0 success, 1 error, 2 unknown failure).  Progress and
 * errors go to stdout. */
int etn_main(int argc, const char* const* argv);

#ifdef __cplusplus
}
#endif
#endif
