// extern "C" entry points of libetn.so (include/etn/etn.h).  Exceptions of the
// C++ layer are caught here and turned into etn_status + etn_last_error().
#include <cctype>
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "batch.hpp"
#include "block_iterative.hpp"
#include "etn/etn.h"
#include "unet.hpp"
#include "wbp.hpp"

namespace {

thread_local std::string g_last_error;

template <typename F>
etn_status guarded(F&& f)
{
    try {
        f();
        return ETN_OK;
    } catch (const etn::Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return ETN_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return ETN_ERROR_INTERNAL;
    } catch (...) {
        g_last_error = "unknown error";
        return ETN_ERROR_INTERNAL;
    }
}

void need(const void* p, const char* what)
{
    if (!p) etn::fail(std::string(what) + " is null");
}

etn::VolumeProperties volume_props(const etn_volume_geometry* v)
{
    need(v, "volume geometry");
    return etn::VolumeProperties(*v);
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

namespace etn {
etn_status record_error(etn_status s, const std::string& message)
{
    g_last_error = message;
    return s;
}
}  // namespace etn

struct etn_reconstruction {
    std::unique_ptr<etn::ReconstructionAlgorithm> op;
};

struct etn_batch {
    std::unique_ptr<etn::BatchReconstruction> batch;
};

struct etn_unet {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::unique_ptr<etn::UNet3D> net;
    ~etn_unet()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        net.reset();
        if (own_stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

const char* etn_last_error(void) { return g_last_error.c_str(); }

const char* etn_version(void) { return "etn 0.1 (gfx950)"; }

etn_status etn_rotate_scanner_geometry(const etn_scanner_geometry* base, float tilt_deg, float x_axis_tilt_deg,
                                       etn_scanner_geometry* out)
{
    return guarded([&] {
        need(base, "base geometry");
        need(out, "output geometry");
        const etn::SingleAxisTiltRotator rot{etn::ParallelScannerGeometry(*base)};
        *out = rot.rotated(tilt_deg, x_axis_tilt_deg).to_c();
    });
}

etn_status etn_ray_direction(const etn_scanner_geometry* g, float out[3])
{
    return guarded([&] {
        need(g, "geometry");
        need(out, "output");
        etn::ParallelScannerGeometry(*g).ray_direction().store(out);
    });
}

etn_status etn_projection_matrix(const etn_scanner_geometry* g, float out[16])
{
    return guarded([&] {
        need(g, "geometry");
        need(out, "output");
        std::memcpy(out, etn::ParallelScannerGeometry(*g).projection_matrix().data(), 16 * sizeof(float));
    });
}

etn_status etn_traversal_length_through_bbox(const etn_scanner_geometry* g, const float bbox_min[3],
                                             const float bbox_max[3], float* out)
{
    return guarded([&] {
        need(g, "geometry");
        need(bbox_min, "bbox_min");
        need(bbox_max, "bbox_max");
        need(out, "output");
        *out = etn::ParallelScannerGeometry(*g).traversal_length_through(etn::Vec3f(bbox_min), etn::Vec3f(bbox_max));
    });
}

etn_status etn_bp_sample_pattern(const etn_scanner_geometry* g, const float voxel_size[3], uint32_t samples,
                                 float* out4)
{
    return guarded([&] {
        need(g, "geometry");
        need(voxel_size, "voxel size");
        need(out4, "output");
        const auto pattern = etn::bp_sample_pattern(etn::ParallelScannerGeometry(*g), etn::Vec3f(voxel_size), samples);
        for (size_t i = 0; i < pattern.size(); ++i) {
            out4[4 * i + 0] = pattern[i].dx;
            out4[4 * i + 1] = pattern[i].dy;
            out4[4 * i + 2] = pattern[i].dz;
            out4[4 * i + 3] = pattern[i].length;
        }
    });
}

etn_status etn_subvolume(const etn_volume_geometry* v, uint32_t index, uint32_t resolution_out[3],
                         uint32_t* first_slice, float bbox_min_out[3], float bbox_max_out[3])
{
    return guarded([&] {
        const etn::VolumeProperties p = volume_props(v);
        const etn::SubVolume& s = p.subvolume(index);
        if (resolution_out) std::memcpy(resolution_out, s.res, sizeof(s.res));
        if (first_slice) *first_slice = s.first_slice;
        if (bbox_min_out) s.bmin.store(bbox_min_out);
        if (bbox_max_out) s.bmax.store(bbox_max_out);
    });
}

etn_status etn_projection_order(const char* scheme, uint32_t count, const uint32_t* skip, uint32_t skip_count,
                                uint32_t* out, uint32_t* out_count)
{
    return guarded([&] {
        need(out_count, "out_count");
        if (count) need(out, "output");
        if (skip_count) need(skip, "skip list");
        auto set = etn::ProjectionIteratorFactory::instantiate(scheme ? scheme : "random");
        std::set<uint32_t> sk(skip, skip + skip_count);
        set->setProjectionCount(count, sk);
        const auto& idx = set->indices();
        std::copy(idx.begin(), idx.end(), out);
        *out_count = (uint32_t)idx.size();
    });
}

etn_status etn_forward_project(const etn_volume_geometry* vol, etn_voxel_type voxel_type, const void* volume,
                               const uint8_t* mask, uint32_t subvolume, const etn_scanner_geometry* g,
                               uint32_t samples, float* projection, float* traversal_length, void* stream)
{
    return guarded([&] {
        need(g, "scanner geometry");
        need(volume, "volume");
        const etn::GpuVolume v(*vol, voxel_type, const_cast<void*>(volume));
        etn::GeometricSetup setup(etn::ParallelScannerGeometry(*g), &v.properties());
        const etn::GpuMask m{mask};
        etn::ForwardProjectionOperator op(&setup, &v, &m, samples, as_stream(stream));
        etn::DeviceImage p{projection, g->resolution[0], g->resolution[1]};
        etn::DeviceImage l{traversal_length, g->resolution[0], g->resolution[1]};
        op.SetOutput(&p, &l);
        op.Execute(subvolume);
    });
}

etn_status etn_back_project(const etn_volume_geometry* vol, etn_voxel_type voxel_type, void* volume,
                            const uint8_t* mask, uint32_t subvolume, const etn_scanner_geometry* g, uint32_t samples,
                            const float* residual, const float* ray_length, float lambda, uint32_t block_size,
                            void* stream)
{
    return guarded([&] {
        need(g, "scanner geometry");
        need(volume, "volume");
        etn::GpuVolume v(*vol, voxel_type, volume);
        etn::GeometricSetup setup(etn::ParallelScannerGeometry(*g), &v.properties());
        const etn::GpuMask m{mask};
        etn::BackProjectionOperator op(&setup, &v, &m, lambda, samples, as_stream(stream));
        const etn::DeviceImage r{const_cast<float*>(residual), g->resolution[0], g->resolution[1]};
        const etn::DeviceImage l{const_cast<float*>(ray_length), g->resolution[0], g->resolution[1]};
        op.setInput(&r, &l, block_size);
        op.execute(subvolume);
    });
}

etn_status etn_compare(uint32_t nu, uint32_t nv, const float* real, const float* virtual_projection, float* residual,
                       void* stream)
{
    return guarded([&] {
        etn::CompareKernel k(as_stream(stream));
        const etn::DeviceImage a{const_cast<float*>(real), nu, nv};
        const etn::DeviceImage b{const_cast<float*>(virtual_projection), nu, nv};
        etn::DeviceImage c{residual, nu, nv};
        if (!real || !virtual_projection || !residual) etn::fail("compare images must not be null");
        k.setInput(&a, &b);
        k.setOutput(&c);
        k.run();
    });
}

etn_status etn_long_object_compensate(uint32_t nu, uint32_t nv, float* projection, float* traversal_length,
                                      float constant_length, void* stream)
{
    return guarded([&] {
        if (!projection || !traversal_length) etn::fail("images must not be null");
        etn::VirtualProjectionKernel k(as_stream(stream));
        etn::DeviceImage p{projection, nu, nv}, l{traversal_length, nu, nv};
        k.setProjection(&p);
        k.setTraversalLength(&l);
        k.setConstantLength(constant_length);
        k.run();
    });
}

etn_status etn_compute_ray_length(const etn_volume_geometry* vol, const etn_scanner_geometry* g, float* ray_length,
                                  void* stream)
{
    return guarded([&] {
        need(g, "scanner geometry");
        const etn::VolumeProperties p = volume_props(vol);
        etn::GeometricSetup setup(etn::ParallelScannerGeometry(*g), &p);
        etn::DeviceImage out{ray_length, g->resolution[0], g->resolution[1]};
        etn::ComputeRayLengthKernel k(&setup, &out, as_stream(stream));
        k.run();
    });
}

etn_status etn_sum_squared_difference_f64(const float* a, const float* b, size_t count, double* out_device,
                                          void* stream)
{
    return guarded([&] {
        need(a, "input");
        need(out_device, "output");
        hipStream_t s = as_stream(stream);
        double* partials = nullptr;
        // stream-ordered scratch of kSsdBlocks doubles, released on the same stream
        ETN_HIP(hipMallocAsync((void**)&partials, sizeof(double) * etn::dev::kSsdBlocks, s));
        etn::dev::launch_sum_squared_difference(a, b, count, partials, out_device, s);
        ETN_HIP(hipFreeAsync(partials, s));
    });
}

etn_status etn_sum_squared_difference_f64_noalloc(const float* a, const float* b, size_t count, double* out_device,
                                                  void* stream)
{
    return guarded([&] {
        need(a, "input");
        need(out_device, "output");
        if (count > ETN_SSD_NOALLOC_MAX_COUNT) etn::fail("allocation-free residual norm: count above the maximum");
        hipStream_t s = as_stream(stream);
        // the scratch of the GPU the stream runs on (the current one for the null stream)
        int device = 0;
        if (s) ETN_HIP(hipStreamGetDevice(s, &device));
        else ETN_HIP(hipGetDevice(&device));
        static_assert(etn::dev::kSsdBlocks * sizeof(double) <= ETN_SSD_NOALLOC_SCRATCH_BYTES,
                      "the partial sums must fit the static scratch");
        etn::dev::launch_sum_squared_difference(a, b, count, etn::dev::ssd_static_scratch(device), out_device, s);
    });
}

etn_status etn_convert_volume(const void* src, etn_voxel_type src_type, void* dst, etn_voxel_type dst_type,
                              size_t count, void* stream)
{
    return guarded([&] {
        need(src, "source");
        need(dst, "destination");
        (void)etn::voxel_bytes(src_type);
        (void)etn::voxel_bytes(dst_type);
        etn::dev::launch_convert(src, src_type, dst, dst_type, count, as_stream(stream));
    });
}

etn_status etn_reconstruction_create(const etn_reconstruction_desc* desc, void* stream, etn_reconstruction** out)
{
    return guarded([&] {
        need(desc, "descriptor");
        need(out, "output handle");
        *out = nullptr;
        auto r = std::make_unique<etn_reconstruction>();
        // Framework::instantiateReconstructionAlgorithm (Framework.cpp:142-152): the
        // block-iterative ids, else the plugin's "wbp" (WBPPlugin.cpp:21-24)
        std::string id = desc->algorithm ? desc->algorithm : "";
        for (char& c : id) c = (char)std::tolower((unsigned char)c);
        if (id == "wbp") r->op = std::make_unique<etn::WeightedBackProjection>(*desc, as_stream(stream));
        else r->op = std::make_unique<etn::BlockIterativeReconstructionOperator>(*desc, as_stream(stream));
        *out = r.release();
    });
}

etn_status etn_reconstruction_destroy(etn_reconstruction* r)
{
    return guarded([&] { delete r; });
}

etn_status etn_reconstruction_run(etn_reconstruction* r)
{
    return guarded([&] {
        need(r, "reconstruction");
        r->op->run();
    });
}

etn_status etn_reconstruction_run_single_iteration(etn_reconstruction* r)
{
    return guarded([&] {
        need(r, "reconstruction");
        r->op->runSingleIteration();
    });
}

etn_status etn_reconstruction_run_iterations(etn_reconstruction* r, uint32_t iterations)
{
    return guarded([&] {
        need(r, "reconstruction");
        r->op->runNiterations(iterations);
    });
}

etn_status etn_reconstruction_volume(etn_reconstruction* r, void** device_volume, size_t* bytes)
{
    return guarded([&] {
        need(r, "reconstruction");
        etn::GpuVolume* v = r->op->getReconstructedVolume();
        if (device_volume) *device_volume = v->data();
        if (bytes) *bytes = v->bytes();
    });
}

etn_status etn_reconstruction_set_volume(etn_reconstruction* r, const void* device_volume)
{
    return guarded([&] {
        need(r, "reconstruction");
        need(device_volume, "source volume");
        etn::GpuVolume* v = r->op->getReconstructedVolume();
        ETN_HIP(hipMemcpyAsync(v->data(), device_volume, v->bytes(), hipMemcpyDeviceToDevice, r->op->stream()));
    });
}

etn_status etn_reconstruction_residual_norm(etn_reconstruction* r, double* out)
{
    return guarded([&] {
        need(r, "reconstruction");
        need(out, "output");
        *out = r->op->residualNorm();
    });
}

etn_status etn_reconstruction_synchronize(etn_reconstruction* r)
{
    return guarded([&] {
        need(r, "reconstruction");
        r->op->synchronize();
    });
}

etn_status etn_reconstruction_uses_fused_path(etn_reconstruction* r, int* out)
{
    return guarded([&] {
        need(r, "reconstruction");
        need(out, "output");
        *out = r->op->usesFusedPath() ? 1 : 0;
    });
}

etn_status etn_reconstruction_time_kernels(etn_reconstruction* r, int* kernel, double* average_ms, uint32_t* launches)
{
    return guarded([&] {
        need(r, "reconstruction");
        need(kernel, "kernel");
        need(average_ms, "average_ms");
        need(launches, "launches");
        *average_ms = r->op->timeKernels(kernel, launches);
    });
}

etn_status etn_reconstruction_set_tile(etn_reconstruction* r, int planes, int band)
{
    return guarded([&] {
        need(r, "reconstruction");
        r->op->setTile(planes, band);
    });
}

etn_status etn_reconstruction_tile(etn_reconstruction* r, int* planes, int* band)
{
    return guarded([&] {
        need(r, "reconstruction");
        need(planes, "planes");
        need(band, "band");
        r->op->ensureTuned();
        r->op->tile(planes, band);
    });
}

etn_status etn_reconstruction_set_workgroup_threads(etn_reconstruction* r, int threads)
{
    return guarded([&] {
        need(r, "reconstruction");
        r->op->setWorkgroupThreads(threads);
    });
}

etn_status etn_reconstruction_workgroup_threads(etn_reconstruction* r, int* threads)
{
    return guarded([&] {
        need(r, "reconstruction");
        need(threads, "threads");
        r->op->ensureTuned();
        *threads = r->op->workgroupThreads();
    });
}

etn_status etn_batch_create(const etn_reconstruction_desc* members, uint32_t count, uint32_t flags, void* stream,
                            etn_batch** out)
{
    return guarded([&] {
        need(members, "member descriptions");
        need(out, "output");
        *out = nullptr;
        auto b = std::make_unique<etn_batch>();
        b->batch = std::make_unique<etn::BatchReconstruction>(members, count, flags, as_stream(stream));
        *out = b.release();
    });
}

etn_status etn_batch_destroy(etn_batch* b)
{
    return guarded([&] { delete b; });
}

etn_status etn_batch_run_iterations(etn_batch* b, uint32_t iterations)
{
    return guarded([&] {
        need(b, "batch");
        b->batch->runNiterations(iterations);
    });
}

etn_status etn_batch_synchronize(etn_batch* b)
{
    return guarded([&] {
        need(b, "batch");
        b->batch->synchronize();
    });
}

etn_status etn_batch_volume(etn_batch* b, uint32_t member, void** device_volume, size_t* bytes)
{
    return guarded([&] {
        need(b, "batch");
        etn::GpuVolume* v = b->batch->member(member).getReconstructedVolume();
        if (device_volume) *device_volume = v->data();
        if (bytes) *bytes = v->bytes();
    });
}

etn_status etn_batch_residual_norm(etn_batch* b, uint32_t member, double* out)
{
    return guarded([&] {
        need(b, "batch");
        need(out, "output");
        *out = b->batch->residualNorm(member);
    });
}

etn_status etn_batch_uses_fused_path(etn_batch* b, uint32_t member, int* out)
{
    return guarded([&] {
        need(b, "batch");
        need(out, "output");
        *out = b->batch->usesFusedPath(member) ? 1 : 0;
    });
}

etn_status etn_batch_tile(etn_batch* b, uint32_t member, int* planes, int* band)
{
    return guarded([&] {
        need(b, "batch");
        need(planes, "planes");
        need(band, "band");
        b->batch->tile(member, planes, band);
    });
}

etn_status etn_batch_workgroup_threads(etn_batch* b, uint32_t member, int* threads)
{
    return guarded([&] {
        need(b, "batch");
        need(threads, "threads");
        *threads = b->batch->workgroupThreads(member);
    });
}

etn_status etn_batch_time_iteration(etn_batch* b, double* fused_ms, uint32_t* launches)
{
    return guarded([&] {
        need(b, "batch");
        need(fused_ms, "fused_ms");
        need(launches, "launches");
        *fused_ms = b->batch->timeIteration(launches);
    });
}

etn_status etn_unet_create(uint32_t nx, uint32_t ny, uint32_t nz, void* stream, etn_unet** out)
{
    return guarded([&] {
        need(out, "output");
        *out = nullptr;
        auto n = std::make_unique<etn_unet>();
        if (stream) {
            n->stream = as_stream(stream);
        } else {
            ETN_HIP(hipStreamCreate(&n->stream));
            n->own_stream = true;
        }
        n->net = std::make_unique<etn::UNet3D>((int)nx, (int)ny, (int)nz, n->stream);
        *out = n.release();
    });
}

etn_status etn_unet_destroy(etn_unet* n)
{
    return guarded([&] { delete n; });
}

etn_status etn_unet_parameter_count(size_t* out)
{
    return guarded([&] {
        need(out, "output");
        *out = etn::UNet3D::parameterCount();
    });
}

etn_status etn_unet_set_parameters(etn_unet* n, const float* host, size_t count)
{
    return guarded([&] {
        need(n, "network");
        need(host, "parameters");
        n->net->setParameters(host, count);
    });
}

etn_status etn_unet_get_parameters(etn_unet* n, float* host, size_t count)
{
    return guarded([&] {
        need(n, "network");
        need(host, "parameters");
        n->net->getParameters(host, count);
    });
}

etn_status etn_unet_predict(etn_unet* n, const float* x, const float* d, float scale, float update_scale,
                            float* out)
{
    return guarded([&] {
        need(n, "network");
        need(x, "input volume");
        need(out, "output volume");
        n->net->predict(x, d, scale, update_scale, out);
    });
}

etn_status etn_unet_gradient(etn_unet* n, const float* x, const float* d, const float* target, float scale,
                             float update_scale, float* grad_host, double* loss)
{
    return guarded([&] {
        need(n, "network");
        need(x, "input volume");
        need(target, "target volume");
        const double l = n->net->gradient(x, d, target, scale, update_scale);
        if (loss) *loss = l;
        if (grad_host) n->net->copyGradient(grad_host, etn::UNet3D::parameterCount());
    });
}

etn_status etn_unet_train_step(etn_unet* n, const float* x, const float* d, const float* target, float scale,
                               float update_scale, float lr, double* loss)
{
    return guarded([&] {
        need(n, "network");
        need(x, "input volume");
        need(target, "target volume");
        const double l = n->net->trainStep(x, d, target, scale, update_scale, lr);
        if (loss) *loss = l;
    });
}

etn_status etn_unet_time_forward(etn_unet* n, const float* x, float scale, double* ms, double* flops)
{
    return guarded([&] {
        need(n, "network");
        need(x, "input volume");
        need(ms, "ms");
        *ms = n->net->timeForward(x, scale, flops);
    });
}

etn_status etn_ramp_filter(uint32_t nu, uint32_t nv, const float* projection, float* filtered, void* stream)
{
    return guarded([&] {
        need(projection, "projection");
        need(filtered, "output");
        if (!nu || !nv) etn::fail("empty image");
        const auto p = etn::WeightedBackProjection::filterParams(nu, nv, etn::WeightedBackProjection::kFilterSize);
        const std::vector<float> k = etn::WeightedBackProjection::ramLakKernel(etn::WeightedBackProjection::kFilterSize);
        etn::DeviceBuffer taps(k.size() * sizeof(float));
        ETN_HIP(hipMemcpyAsync(taps.get(), k.data(), taps.bytes(), hipMemcpyHostToDevice, as_stream(stream)));
        etn::dev::launch_ramp_filter(p, taps.as<float>(), projection, filtered, as_stream(stream));
        ETN_HIP(hipStreamSynchronize(as_stream(stream)));   // taps are freed on return
    });
}

}  // extern "C"
