#include "wbp.hpp"

#include <algorithm>

#include "geometry.hpp"

namespace etn {

namespace {

// NumericalAlgorithms::nearestPowerOfTwo (NumericalAlgorithms.cpp:20-31)
uint32_t next_power_of_two(uint32_t v)
{
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}

}  // namespace

std::vector<float> WeightedBackProjection::ramLakKernel(uint32_t resolution)
{
    const float pi = 3.14159265358979323846f;
    const float pi2 = pi * pi;   // M_PI_F * M_PI_F, a float product
    std::vector<float> k(resolution);
    for (uint32_t i = 0; i < resolution; ++i) {
        const int position = (int)i - (int)(resolution / 2);
        if (position == 0) k[i] = 1.0f / 4.0f;
        else if (position % 2 == 0) k[i] = 0.0f;
        else k[i] = -1.0f / ((float)(position * position) * pi2);
    }
    return k;
}

dev::RampFilterParams WeightedBackProjection::filterParams(uint32_t nu, uint32_t nv, uint32_t taps)
{
    if (taps == 0 || taps % 2) fail("the ramp filter needs an even, non-zero number of taps");
    dev::RampFilterParams p{};
    p.nu = nu;
    p.nv = nv;
    // ConvolutionFourierSpaceImplementation::getUnifiedResolution: the larger
    // of filter (taps x 1) and image, rounded up to powers of two
    p.ux = next_power_of_two(std::max(taps, nu));
    p.uy = next_power_of_two(std::max(1u, nv));
    p.pad_x = ((int32_t)p.ux - (int32_t)nu) / 2;
    p.pad_y = ((int32_t)p.uy - (int32_t)nv) / 2;
    p.end_x = ((int32_t)p.ux + (int32_t)nu) / 2;
    p.end_y = ((int32_t)p.uy + (int32_t)nv) / 2;
    p.crop_x = (int32_t)(p.ux / 2) - (int32_t)(nu / 2);
    p.crop_y = (int32_t)(p.uy / 2) - (int32_t)(nv / 2);
    p.taps = taps;
    if (p.ux > dev::kMaxRampRow)
        unsupported("ramp filter: detector rows longer than " + std::to_string(dev::kMaxRampRow / 2) +
                    " pixels are not supported");
    return p;
}

WeightedBackProjection::WeightedBackProjection(const etn_reconstruction_desc& d, hipStream_t stream)
    : count_(d.projection_count), lambda_(d.lambda), projections_(static_cast<const float*>(d.projections))
{
    if (d.projection_count == 0 || (!d.tilt_angles && !d.geometries)) fail("No projections given!");
    if (!d.projections) fail("projection stack device pointer is null");
    if (d.flags & ETN_RECON_FLAG_16BIT_PROJECTIONS) unsupported("WBP filters f32 projections");
    const uint32_t nu = d.base_geometry.resolution[0], nv = d.base_geometry.resolution[1];
    if (nu == 0 || nv == 0) fail("detector resolution is zero");
    pixels_ = (size_t)nu * nv;
    params_ = filterParams(nu, nv, kFilterSize);

    geometries_ = projection_geometries(d);

    if (stream) {
        stream_ = stream;
    } else {
        ETN_HIP(hipStreamCreate(&stream_));
        own_stream_ = true;
    }
    etn_volume_geometry vg = d.volume;
    if (vg.subvolume_count == 0) vg.subvolume_count = 1;
    volume_ = std::make_unique<GpuVolume>(vg, d.voxel_type);
    volume_->fill(d.init_value, stream_);
    setup_ = std::make_unique<GeometricSetup>(geometries_[0], &volume_->properties());
    backward_ = std::make_unique<BackProjectionOperator>(setup_.get(), volume_.get(), &mask_, lambda_,
                                                         d.bp_samples ? d.bp_samples : 1, stream_);
    images_.allocate(2 * pixels_ * sizeof(float));
    filtered_ = {images_.as<float>(), nu, nv};
    length_ = {images_.as<float>() + pixels_, nu, nv};
    ray_length_ = std::make_unique<ComputeRayLengthKernel>(setup_.get(), &length_, stream_);
    const std::vector<float> k = ramLakKernel(kFilterSize);
    taps_.allocate(k.size() * sizeof(float));
    ETN_HIP(hipMemcpyAsync(taps_.get(), k.data(), taps_.bytes(), hipMemcpyHostToDevice, stream_));
    ETN_HIP(hipStreamSynchronize(stream_));
    if ((d.flags & (ETN_RECON_FLAG_GENERIC_KERNELS | ETN_RECON_FLAG_EXACT_ARITHMETIC)) == 0)
        fused_ = FusedBackProjector::create(volume_.get(), geometries_, nu, nv, d.bp_samples ? d.bp_samples : 1,
                                            (d.flags & ETN_RECON_FLAG_LDS_BACK_PROJECTION) != 0, stream_);
    if (fused_) {
        std::vector<dev::RayParams> geo;
        for (const ParallelScannerGeometry& g : geometries_) geo.push_back(make_ray_params(g, 1));
        rays_.allocate(sizeof(dev::RayParams) * std::max<size_t>(geo.size(), 1));
        ETN_HIP(hipMemcpyAsync(rays_.get(), geo.data(), sizeof(dev::RayParams) * geo.size(), hipMemcpyHostToDevice,
                               stream_));
        ETN_HIP(hipStreamSynchronize(stream_));
    }
}

WeightedBackProjection::~WeightedBackProjection()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (own_stream_) (void)hipStreamDestroy(stream_);
}

void WeightedBackProjection::filter(const float* src, float* dst)
{
    dev::launch_ramp_filter(params_, taps_.as<float>(), src, dst, stream_);
}

void WeightedBackProjection::run() { runNiterations(count_); }

void WeightedBackProjection::runNiterations(uint32_t n)
{
    if (n > count_) fail("WBP: more projections requested than the stack holds");
    if (fused_) {
        // filtered images and ray lengths of up to 64 projections, then their
        // back projections in stack order, eight per volume sweep
        constexpr uint32_t kChunk = 64;
        DeviceBuffer chunk(sizeof(float) * 2 * pixels_ * std::min(n, kChunk));
        const VolumeProperties& vp = setup_->getVolume();
        float bmin[3], bmax[3];
        vp.bbox_min().store(bmin);
        vp.bbox_max().store(bmax);
        for (uint32_t i0 = 0; i0 < n; i0 += kChunk) {
            std::vector<uint32_t> idx;
            std::vector<const float*> res, len;
            const uint32_t i1 = std::min(n, i0 + kChunk);
            // the chunk's projections filtered in one launch and their ray
            // lengths in another (one launch per image had cost 13 and 4 us
            // each, most of it launch and ramp-up)
            dev::launch_ramp_filter_batch(params_, taps_.as<float>(), projections_ + (size_t)i0 * pixels_, pixels_,
                                          chunk.as<float>(), 2 * pixels_, i1 - i0, stream_);
            dev::launch_ray_length_batch(bmin, bmax, rays_.as<dev::RayParams>() + i0, i1 - i0,
                                         geometries_[i0].nu(), geometries_[i0].nv(),
                                         chunk.as<float>() + pixels_, 2 * pixels_, stream_);
            for (uint32_t i = i0; i < i1; ++i) {
                float* f = chunk.as<float>() + 2 * pixels_ * (i - i0);
                idx.push_back(i);
                res.push_back(f);
                len.push_back(f + pixels_);
            }
            // backProject synchronises the stream before it returns, so the
            // next chunk may reuse the buffer
            fused_->backProject(idx, res, len, lambda_ / (float)count_, stream_, launch_ms_);
        }
        return;
    }
    for (uint32_t i = 0; i < n; ++i) {
        setup_->setScannerGeometry(geometries_[i]);
        filter(projections_ + (size_t)i * pixels_, filtered_.data);
        ray_length_->run();
        backward_->setInput(&filtered_, &length_, count_);
        backward_->execute();
        backward_->setInput(nullptr, nullptr, 1);
    }
}

void WeightedBackProjection::synchronize() { ETN_HIP(hipStreamSynchronize(stream_)); }

double WeightedBackProjection::timeKernels(int* kernel, uint32_t* launches)
{
    std::vector<float> ms;
    if (fused_) {
        *kernel = 0;
        launch_ms_ = &ms;
        try {
            run();
        } catch (...) {
            launch_ms_ = nullptr;
            throw;
        }
        launch_ms_ = nullptr;
    } else {
        *kernel = 2;
        hipEvent_t a, b;
        ETN_HIP(hipEventCreate(&a));
        ETN_HIP(hipEventCreate(&b));
        for (uint32_t i = 0; i < count_; ++i) {
            setup_->setScannerGeometry(geometries_[i]);
            filter(projections_ + (size_t)i * pixels_, filtered_.data);
            ray_length_->run();
            backward_->setInput(&filtered_, &length_, count_);
            ETN_HIP(hipEventRecord(a, stream_));
            backward_->execute();
            ETN_HIP(hipEventRecord(b, stream_));
            backward_->setInput(nullptr, nullptr, 1);
            ETN_HIP(hipEventSynchronize(b));
            float t = 0.0f;
            ETN_HIP(hipEventElapsedTime(&t, a, b));
            ms.push_back(t);
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    *launches = (uint32_t)ms.size();
    double total = 0.0;
    for (float t : ms) total += t;
    return ms.empty() ? 0.0 : total / ms.size();
}

}  // namespace etn
