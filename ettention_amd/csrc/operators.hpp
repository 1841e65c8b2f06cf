// C++ operator surface of the hot path, named after the reference's classes:
// GeometricSetup, the forward/back projection kernels, operators and their
// creator factories, CompareKernel, VirtualProjectionKernel and
// ComputeRayLengthKernel.  Device work is stream-ordered on the HIP stream
// the object is constructed with.
#pragma once

#include <memory>
#include <vector>

#include "geometry.hpp"
#include "kernels.hpp"

namespace etn {

// GPU-resident volume (the device half of the reference's GPUMappedVolume):
// properties, storage type and the complete volume in linear z-major layout.
class GpuVolume {
public:
    GpuVolume(const etn_volume_geometry& g, etn_voxel_type t);                   // owns its storage
    GpuVolume(const etn_volume_geometry& g, etn_voxel_type t, void* device);      // borrows device memory
    const VolumeProperties& properties() const { return props_; }
    etn_voxel_type voxel_type() const { return type_; }
    void* data() const { return data_; }
    size_t bytes() const { return props_.voxel_count() * voxel_bytes(type_); }
    void fill(float value, hipStream_t stream);

private:
    VolumeProperties props_;
    etn_voxel_type type_;
    DeviceBuffer owned_;
    void* data_ = nullptr;
};

// Prior-knowledge mask: one uint8 per voxel of the complete volume (0..255).
struct GpuMask {
    const uint8_t* data = nullptr;
};

// Device detector image (the device half of GPUMapped<Image>).
struct DeviceImage {
    float* data = nullptr;
    uint32_t nu = 0, nv = 0;
    size_t pixels() const { return (size_t)nu * nv; }
};

// GeometricSetup (GeometricSetup.cpp:53-112): current scanner geometry and the
// volume whose subvolume the next kernel works on.
class GeometricSetup {
public:
    GeometricSetup(const ParallelScannerGeometry& g, const VolumeProperties* volume) : geometry_(g), volume_(volume) {}
    void setScannerGeometry(const ParallelScannerGeometry& g) { geometry_ = g; }
    const ParallelScannerGeometry& getScannerGeometry() const { return geometry_; }
    const VolumeProperties& getVolume() const { return *volume_; }
    void setVolume(const VolumeProperties* v) { volume_ = v; }
    void setCurrentSubVolumeIndex(uint32_t i) { subvolume_ = i; }
    uint32_t getCurrentSubVolumeIndex() const { return subvolume_; }

private:
    ParallelScannerGeometry geometry_;
    const VolumeProperties* volume_;
    uint32_t subvolume_ = 0;
};

dev::SlabParams make_slab_params(const VolumeProperties& v, uint32_t subvolume);
dev::RayParams make_ray_params(const ParallelScannerGeometry& g, uint32_t samples);
dev::BpParams make_bp_params(const ParallelScannerGeometry& g, const VolumeProperties& v, uint32_t samples,
                             float lambda_over_block);

// ForwardProjectionKernel (ForwardProjectionKernel.cpp:11-118).
class ForwardProjectionKernel {
public:
    virtual ~ForwardProjectionKernel() = default;
    // both or neither, equal sizes (ForwardProjectionKernel.cpp:36-57)
    void setOutput(DeviceImage* projection, DeviceImage* traversal_length);
    virtual void run() = 0;
    virtual const char* getKernelName() const = 0;

protected:
    DeviceImage* projection_ = nullptr;
    DeviceImage* traversal_length_ = nullptr;
};

// ParallelBeamsForwardProjectionKernel: the generic DDA kernel of projectors.hip.
class ParallelBeamsForwardProjectionKernel final : public ForwardProjectionKernel {
public:
    ParallelBeamsForwardProjectionKernel(const GeometricSetup* setup, const GpuVolume* volume, const GpuMask* mask,
                                         uint32_t samples, hipStream_t stream);
    void run() override;
    const char* getKernelName() const override { return "ParallelBeamsForwardProjectionKernel (gfx950)"; }

private:
    const GeometricSetup* setup_;
    const GpuVolume* volume_;
    const GpuMask* mask_;
    uint32_t samples_;
    hipStream_t stream_;
};

// ForwardProjectionFactory (factory/ForwardProjectionFactory.cpp:16-42): creators
// are asked newest first; a null result means "not mine".
class ForwardProjectionFactory {
public:
    struct InstanceCreator {
        virtual ~InstanceCreator() = default;
        virtual std::unique_ptr<ForwardProjectionKernel> CreateKernelInstance(const GeometricSetup* setup,
                                                                              const GpuVolume* volume,
                                                                              const GpuMask* mask, uint32_t samples,
                                                                              hipStream_t stream) const = 0;
    };
    static ForwardProjectionFactory& instance();
    void RegisterCreator(std::unique_ptr<InstanceCreator> creator);
    std::unique_ptr<ForwardProjectionKernel> CreateKernelInstance(const GeometricSetup* setup, const GpuVolume* volume,
                                                                  const GpuMask* mask, uint32_t samples,
                                                                  hipStream_t stream) const;

private:
    ForwardProjectionFactory();
    std::vector<std::unique_ptr<InstanceCreator>> creators_;
};

// ForwardProjectionOperator (ForwardProjectionOperator.cpp:11-55): Execute adds
// the subvolume's line integrals into the outputs.
class ForwardProjectionOperator {
public:
    ForwardProjectionOperator(GeometricSetup* setup, const GpuVolume* volume, const GpuMask* mask, uint32_t samples,
                              hipStream_t stream);
    void SetOutput(DeviceImage* projection, DeviceImage* traversal_length);
    void Execute(uint32_t subvolume);
    ForwardProjectionKernel* GetKernel() const { return kernel_.get(); }

private:
    GeometricSetup* setup_;
    std::unique_ptr<ForwardProjectionKernel> kernel_;
};

// BackProjectionKernel (BackProjectionKernel.cpp:14-184).
class BackProjectionKernel {
public:
    virtual ~BackProjectionKernel() = default;
    // both or neither, equal sizes (BackProjectionKernel.cpp:127-144)
    void SetInput(const DeviceImage* residual, const DeviceImage* ray_length);
    void SetProjectionBlockSize(uint32_t b) { block_size_ = b ? b : 1; }
    virtual void run() = 0;
    virtual const char* getKernelName() const = 0;

protected:
    const DeviceImage* residual_ = nullptr;
    const DeviceImage* ray_length_ = nullptr;
    uint32_t block_size_ = 1;
};

// ParallelBeamsBackProjectionKernel (ParallelBeamsBackProjectionKernel.cpp:13-103).
class ParallelBeamsBackProjectionKernel final : public BackProjectionKernel {
public:
    ParallelBeamsBackProjectionKernel(const GeometricSetup* setup, GpuVolume* volume, const GpuMask* mask,
                                      float lambda, uint32_t samples, hipStream_t stream);
    void setSamples(uint32_t samples);
    std::vector<VoxelSample> generateSamplingPatternForPerVoxelSampling() const;
    void run() override;
    const char* getKernelName() const override { return "ParallelBeamsBackProjectionKernel (gfx950)"; }

private:
    const GeometricSetup* setup_;
    GpuVolume* volume_;
    const GpuMask* mask_;
    float lambda_;
    uint32_t samples_ = 1;
    hipStream_t stream_;
};

class BackProjectionFactory {
public:
    struct InstanceCreator {
        virtual ~InstanceCreator() = default;
        virtual std::unique_ptr<BackProjectionKernel> CreateKernelInstance(const GeometricSetup* setup,
                                                                           GpuVolume* volume, const GpuMask* mask,
                                                                           float lambda, uint32_t samples,
                                                                           hipStream_t stream) const = 0;
    };
    static BackProjectionFactory& instance();
    void RegisterCreator(std::unique_ptr<InstanceCreator> creator);
    std::unique_ptr<BackProjectionKernel> CreateKernelInstance(const GeometricSetup* setup, GpuVolume* volume,
                                                               const GpuMask* mask, float lambda, uint32_t samples,
                                                               hipStream_t stream) const;

private:
    BackProjectionFactory();
    std::vector<std::unique_ptr<InstanceCreator>> creators_;
};

// BackProjectionOperator (BackProjectionOperator.cpp:13-98).
class BackProjectionOperator {
public:
    static constexpr uint32_t ALL_SUB_VOLUMES = ETN_ALL_SUBVOLUMES;
    BackProjectionOperator(GeometricSetup* setup, GpuVolume* volume, const GpuMask* mask, float lambda,
                           uint32_t samples, hipStream_t stream);
    void setInput(const DeviceImage* residual, const DeviceImage* ray_length, uint32_t block_size);
    void execute(uint32_t subvolume = ALL_SUB_VOLUMES);
    BackProjectionKernel* getKernel() const { return kernel_.get(); }

private:
    GeometricSetup* setup_;
    GpuVolume* volume_;
    std::unique_ptr<BackProjectionKernel> kernel_;
};

// CompareKernel (CompareKernel.cpp:9-66): residual = real - virtual.
class CompareKernel {
public:
    explicit CompareKernel(hipStream_t stream) : stream_(stream) {}
    // real_type: storage of the measured image (f16/bf16 projections; its data
    // pointer then addresses 16-bit values)
    void setInput(const DeviceImage* real, const DeviceImage* virt, etn_voxel_type real_type = ETN_VOXEL_FLOAT32)
    {
        real_ = real;
        virt_ = virt;
        real_type_ = real_type;
    }
    void setOutput(DeviceImage* residual) { residual_ = residual; }
    void run();

private:
    etn_voxel_type real_type_ = ETN_VOXEL_FLOAT32;
    const DeviceImage* real_ = nullptr;
    const DeviceImage* virt_ = nullptr;
    DeviceImage* residual_ = nullptr;
    hipStream_t stream_;
};

// VirtualProjectionKernel / LongObjectCompensationKernel (LongObjectCompensationKernel.cpp:9-59).
class VirtualProjectionKernel {
public:
    explicit VirtualProjectionKernel(hipStream_t stream) : stream_(stream) {}
    void setProjection(DeviceImage* p) { projection_ = p; }
    void setTraversalLength(DeviceImage* l) { length_ = l; }
    void setConstantLength(float c) { c_ = c; }
    void run();

private:
    DeviceImage* projection_ = nullptr;
    DeviceImage* length_ = nullptr;
    float c_ = 0.0f;
    hipStream_t stream_;
};

// ComputeRayLengthKernel (ComputeRayLengthKernel.cpp:9-46).
class ComputeRayLengthKernel {
public:
    ComputeRayLengthKernel(const GeometricSetup* setup, DeviceImage* output, hipStream_t stream)
        : setup_(setup), output_(output), stream_(stream) {}
    void run();
    DeviceImage* getOutput() const { return output_; }

private:
    const GeometricSetup* setup_;
    DeviceImage* output_;
    hipStream_t stream_;
};

}  // namespace etn
