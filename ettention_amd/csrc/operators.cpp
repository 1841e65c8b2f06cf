#include "operators.hpp"

#include <cstring>

namespace etn {

GpuVolume::GpuVolume(const etn_volume_geometry& g, etn_voxel_type t)
    : props_(g), type_(t), owned_(props_.voxel_count() * voxel_bytes(t)), data_(owned_.get())
{
}

GpuVolume::GpuVolume(const etn_volume_geometry& g, etn_voxel_type t, void* device)
    : props_(g), type_(t), data_(device)
{
    (void)voxel_bytes(t);
    if (!device) fail("volume device pointer is null");
}

void GpuVolume::fill(float value, hipStream_t stream) { dev::launch_fill(data_, type_, value, props_.voxel_count(), stream); }

dev::SlabParams make_slab_params(const VolumeProperties& v, uint32_t subvolume)
{
    const SubVolume& s = v.subvolume(subvolume);
    dev::SlabParams p;
    s.bmin.store(p.bmin);
    s.bmax.store(p.bmax);
    v.voxel_size().store(p.vs);
    for (int k = 0; k < 3; ++k) p.res[k] = (int)s.res[k];
    p.nx = v.nx();
    p.ny = v.ny();
    p.first_voxel = s.first_voxel;
    return p;
}

dev::RayParams make_ray_params(const ParallelScannerGeometry& g, uint32_t samples)
{
    dev::RayParams r;
    g.source_base().store(r.source);
    g.horizontal_pitch().store(r.h);
    g.vertical_pitch().store(r.v);
    g.ray_direction().store(r.ray);
    r.nu = g.nu();
    r.nv = g.nv();
    r.samples = samples ? samples : 1;   // ForwardProjectionKernel.cpp:65-68
    return r;
}

dev::BpParams make_bp_params(const ParallelScannerGeometry& g, const VolumeProperties& v, uint32_t samples,
                             float lambda_over_block)
{
    dev::BpParams b;
    std::memcpy(b.m, g.projection_matrix().data(), sizeof(b.m));
    b.lambda = lambda_over_block;
    b.nu = g.nu();
    b.nv = g.nv();
    const std::vector<VoxelSample> pattern = bp_sample_pattern(g, v.voxel_size(), samples);
    if (pattern.size() > (size_t)dev::kMaxBpSamples) unsupported("more than 64 back projection samples");
    b.nsamples = (uint32_t)pattern.size();
    for (size_t i = 0; i < pattern.size(); ++i)
        b.samples[i] = make_float4(pattern[i].dx, pattern[i].dy, pattern[i].dz, pattern[i].length);
    return b;
}

// ---- forward projection ----------------------------------------------------

void ForwardProjectionKernel::setOutput(DeviceImage* projection, DeviceImage* traversal_length)
{
    if (!projection != !traversal_length)
        fail("Virtual projection and traversal length must either be set both or none of them!");
    if (projection && (projection->nu != traversal_length->nu || projection->nv != traversal_length->nv))
        fail("Virtual projection and traversal length images differ in size!");
    projection_ = projection;
    traversal_length_ = traversal_length;
}

ParallelBeamsForwardProjectionKernel::ParallelBeamsForwardProjectionKernel(const GeometricSetup* setup,
                                                                           const GpuVolume* volume,
                                                                           const GpuMask* mask, uint32_t samples,
                                                                           hipStream_t stream)
    : setup_(setup), volume_(volume), mask_(mask), samples_(samples), stream_(stream)
{
}

void ParallelBeamsForwardProjectionKernel::run()
{
    if (!projection_ || !traversal_length_ || !projection_->data || !traversal_length_->data)
        fail("ForwardProjectionKernel cannot be run without setting output images! Be sure to call setOutput() "
             "before with non-null parameters.");
    const ParallelScannerGeometry& g = setup_->getScannerGeometry();
    if (projection_->nu != g.nu() || projection_->nv != g.nv()) fail("output image does not match the detector");
    const dev::SlabParams s = make_slab_params(volume_->properties(), setup_->getCurrentSubVolumeIndex());
    const dev::RayParams r = make_ray_params(g, samples_);
    dev::launch_forward_project(s, r, volume_->voxel_type(), volume_->data(), mask_ ? mask_->data : nullptr,
                                projection_->data, traversal_length_->data, stream_);
}

namespace {
// DefaultForwardProjectionCreator (factory/DefaultForwardProjectionCreator.cpp:28-63):
// parallel geometry with voxel basis -> ParallelBeamsForwardProjectionKernel.
struct DefaultForwardProjectionCreator final : ForwardProjectionFactory::InstanceCreator {
    std::unique_ptr<ForwardProjectionKernel> CreateKernelInstance(const GeometricSetup* setup, const GpuVolume* volume,
                                                                  const GpuMask* mask, uint32_t samples,
                                                                  hipStream_t stream) const override
    {
        return std::make_unique<ParallelBeamsForwardProjectionKernel>(setup, volume, mask, samples, stream);
    }
};
struct DefaultBackProjectionCreator final : BackProjectionFactory::InstanceCreator {
    std::unique_ptr<BackProjectionKernel> CreateKernelInstance(const GeometricSetup* setup, GpuVolume* volume,
                                                               const GpuMask* mask, float lambda, uint32_t samples,
                                                               hipStream_t stream) const override
    {
        return std::make_unique<ParallelBeamsBackProjectionKernel>(setup, volume, mask, lambda, samples, stream);
    }
};
}  // namespace

ForwardProjectionFactory::ForwardProjectionFactory()
{
    RegisterCreator(std::make_unique<DefaultForwardProjectionCreator>());
}

ForwardProjectionFactory& ForwardProjectionFactory::instance()
{
    static ForwardProjectionFactory f;
    return f;
}

void ForwardProjectionFactory::RegisterCreator(std::unique_ptr<InstanceCreator> creator)
{
    creators_.push_back(std::move(creator));
}

std::unique_ptr<ForwardProjectionKernel> ForwardProjectionFactory::CreateKernelInstance(const GeometricSetup* setup,
                                                                                        const GpuVolume* volume,
                                                                                        const GpuMask* mask,
                                                                                        uint32_t samples,
                                                                                        hipStream_t stream) const
{
    for (auto it = creators_.rbegin(); it != creators_.rend(); ++it)
        if (auto k = (*it)->CreateKernelInstance(setup, volume, mask, samples, stream)) return k;
    return nullptr;
}

ForwardProjectionOperator::ForwardProjectionOperator(GeometricSetup* setup, const GpuVolume* volume,
                                                     const GpuMask* mask, uint32_t samples, hipStream_t stream)
    : setup_(setup), kernel_(ForwardProjectionFactory::instance().CreateKernelInstance(setup, volume, mask, samples, stream))
{
    if (!kernel_) unsupported("Parameter set not compatible with known forward projection variants!");
}

void ForwardProjectionOperator::SetOutput(DeviceImage* projection, DeviceImage* traversal_length)
{
    kernel_->setOutput(projection, traversal_length);
}

void ForwardProjectionOperator::Execute(uint32_t subvolume)
{
    (void)setup_->getVolume().subvolume(subvolume);   // range check
    setup_->setCurrentSubVolumeIndex(subvolume);
    kernel_->run();
}

// ---- back projection -------------------------------------------------------

void BackProjectionKernel::SetInput(const DeviceImage* residual, const DeviceImage* ray_length)
{
    if (!residual != !ray_length)
        fail("Either residual and traversalLength images must be provided both or none of them!");
    if (residual && (residual->nu != ray_length->nu || residual->nv != ray_length->nv))
        fail("Residual and traversalLength differ in size!");
    residual_ = residual;
    ray_length_ = ray_length;
}

ParallelBeamsBackProjectionKernel::ParallelBeamsBackProjectionKernel(const GeometricSetup* setup, GpuVolume* volume,
                                                                     const GpuMask* mask, float lambda,
                                                                     uint32_t samples, hipStream_t stream)
    : setup_(setup), volume_(volume), mask_(mask), lambda_(lambda), stream_(stream)
{
    setSamples(samples);
}

void ParallelBeamsBackProjectionKernel::setSamples(uint32_t samples)
{
    if (samples == 0) fail("zero samples not supported for BackProjection");
    const uint32_t m = (uint32_t)std::sqrt((float)samples);
    if (m * m != samples) fail("non-square root number of samples not supported for BackProjection");
    if (samples > (uint32_t)dev::kMaxBpSamples) unsupported("more than 64 back projection samples");
    samples_ = samples;
}

std::vector<VoxelSample> ParallelBeamsBackProjectionKernel::generateSamplingPatternForPerVoxelSampling() const
{
    return bp_sample_pattern(setup_->getScannerGeometry(), volume_->properties().voxel_size(), samples_);
}

void ParallelBeamsBackProjectionKernel::run()
{
    if (!residual_ || !ray_length_ || !residual_->data || !ray_length_->data)
        fail("Residual and traversalLength images must both be provided via SetInput() before kernel launch!");
    const ParallelScannerGeometry& g = setup_->getScannerGeometry();
    if (residual_->nu != g.nu() || residual_->nv != g.nv()) fail("input image does not match the detector");
    const dev::SlabParams s = make_slab_params(volume_->properties(), setup_->getCurrentSubVolumeIndex());
    // lambda / projectionBlockSize at bind time (BackProjectionKernel.cpp:94)
    const dev::BpParams b = make_bp_params(g, volume_->properties(), samples_, lambda_ / (float)block_size_);
    dev::launch_back_project(s, b, volume_->voxel_type(), volume_->data(), mask_ ? mask_->data : nullptr,
                             residual_->data, ray_length_->data, stream_);
}

BackProjectionFactory::BackProjectionFactory() { RegisterCreator(std::make_unique<DefaultBackProjectionCreator>()); }

BackProjectionFactory& BackProjectionFactory::instance()
{
    static BackProjectionFactory f;
    return f;
}

void BackProjectionFactory::RegisterCreator(std::unique_ptr<InstanceCreator> creator)
{
    creators_.push_back(std::move(creator));
}

std::unique_ptr<BackProjectionKernel> BackProjectionFactory::CreateKernelInstance(const GeometricSetup* setup,
                                                                                  GpuVolume* volume,
                                                                                  const GpuMask* mask, float lambda,
                                                                                  uint32_t samples,
                                                                                  hipStream_t stream) const
{
    for (auto it = creators_.rbegin(); it != creators_.rend(); ++it)
        if (auto k = (*it)->CreateKernelInstance(setup, volume, mask, lambda, samples, stream)) return k;
    return nullptr;
}

BackProjectionOperator::BackProjectionOperator(GeometricSetup* setup, GpuVolume* volume, const GpuMask* mask,
                                               float lambda, uint32_t samples, hipStream_t stream)
    : setup_(setup), volume_(volume),
      kernel_(BackProjectionFactory::instance().CreateKernelInstance(setup, volume, mask, lambda, samples, stream))
{
    if (!kernel_) unsupported("No fitting back projection kernel found!");
}

void BackProjectionOperator::setInput(const DeviceImage* residual, const DeviceImage* ray_length, uint32_t block_size)
{
    kernel_->SetInput(residual, ray_length);
    kernel_->SetProjectionBlockSize(block_size);
}

void BackProjectionOperator::execute(uint32_t subvolume)
{
    const uint32_t count = volume_->properties().subvolume_count();
    if (subvolume == ALL_SUB_VOLUMES) {
        for (uint32_t i = 0; i < count; ++i) {
            setup_->setCurrentSubVolumeIndex(i);
            kernel_->run();
        }
        return;
    }
    (void)volume_->properties().subvolume(subvolume);
    setup_->setCurrentSubVolumeIndex(subvolume);
    kernel_->run();
}

// ---- residual, long-object compensation, ray length ------------------------

void CompareKernel::run()
{
    if (!real_ || !virt_ || !residual_) fail("CompareKernel needs input and output images");
    if (real_->pixels() != virt_->pixels() || real_->pixels() != residual_->pixels())
        fail("CompareKernel images differ in size");
    dev::launch_compare((uint32_t)residual_->pixels(), real_->data, real_type_, virt_->data, residual_->data, stream_);
}

void VirtualProjectionKernel::run()
{
    if (!projection_ || !length_) fail("VirtualProjectionKernel needs a projection and a traversal length");
    if (projection_->pixels() != length_->pixels()) fail("projection and traversal length differ in size");
    dev::launch_long_object_compensate((uint32_t)projection_->pixels(), projection_->data, length_->data, c_, stream_);
}

void ComputeRayLengthKernel::run()
{
    if (!output_ || !output_->data) fail("ComputeRayLengthKernel needs an output image");
    const ParallelScannerGeometry& g = setup_->getScannerGeometry();
    const VolumeProperties& v = setup_->getVolume();
    float bmin[3], bmax[3];
    v.bbox_min().store(bmin);
    v.bbox_max().store(bmax);
    dev::launch_ray_length(bmin, bmax, make_ray_params(g, 1), output_->data, stream_);
}

}  // namespace etn
