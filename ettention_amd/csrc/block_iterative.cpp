#include "block_iterative.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>

#include "fused.hpp"

namespace etn {

namespace {
std::string lower(const char* s)
{
    std::string out = s ? s : "";
    for (char& c : out) c = (char)std::tolower((unsigned char)c);
    return out;
}
}  // namespace

BlockIterativeReconstructionOperator::BlockIterativeReconstructionOperator(const etn_reconstruction_desc& d,
                                                                           hipStream_t stream)
    : desc_(d), iterations_(d.iterations), lambda_(d.lambda), use_loc_(d.use_long_object_compensation != 0),
      fp_samples_(d.fp_samples ? d.fp_samples : 1), bp_samples_(d.bp_samples ? d.bp_samples : 1),
      projections_(d.projections), use_graph_((d.flags & ETN_RECON_FLAG_NO_GRAPH) == 0)
{
    // AlgorithmParameterSet::AlgorithmParameterSet (AlgorithmParameterSet.cpp:8-24)
    const std::string algorithm = lower(d.algorithm);
    if (algorithm == "sart") general_block_size_ = 1u;
    else if (algorithm == "sirt") general_block_size_ = 0xffffffffu;
    else if (algorithm == "blockiterative") general_block_size_ = d.block_size;
    else unsupported("unknown reconstruction algorithm '" + algorithm + "'");
    if (general_block_size_ == 0) fail("Illegal ProjectionBlockSize!");
    if (d.projection_count == 0 || (!d.tilt_angles && !d.geometries)) fail("No projections given!");
    if (!d.projections) fail("projection stack device pointer is null");
    if (d.flags & ETN_RECON_FLAG_16BIT_PROJECTIONS) {
        // the measured projections are stored in the volume's 16-bit type
        if (d.voxel_type != ETN_VOXEL_FLOAT16 && d.voxel_type != ETN_VOXEL_BFLOAT16)
            fail("16-bit projections need f16 or bf16 voxel storage (they take the volume's type)");
        projection_type_ = d.voxel_type;
    }
    if (d.base_geometry.resolution[0] == 0 || d.base_geometry.resolution[1] == 0) fail("detector resolution is zero");
    pixels_ = (size_t)d.base_geometry.resolution[0] * d.base_geometry.resolution[1];
    desc_.algorithm = nullptr;
    desc_.projection_iteration = nullptr;
    desc_.skip_projections = nullptr;
    desc_.tilt_angles = nullptr;
    desc_.geometries = nullptr;

    // ProjectionIteratorFactory::instantiate + setProjectionCount (BlockIterativeProjectionAccess.cpp:27-46)
    projection_set_ = ProjectionIteratorFactory::instantiate(d.projection_iteration ? d.projection_iteration : "random");
    std::set<uint32_t> skip;
    for (uint32_t i = 0; i < d.skip_count; ++i) skip.insert(d.skip_projections[i]);
    projection_set_->setProjectionCount(d.projection_count, skip);
    if (projection_set_->getNumberOfProjections() == 0) fail("After applying SkipProjectionList no projections are left!");
    for (uint32_t i : projection_set_->indices())
        if (i >= d.projection_count) fail("projection index out of range");
    block_count_ = (uint32_t)std::ceil((float)projection_set_->getNumberOfProjections() / (float)general_block_size_);

    geometries_ = projection_geometries(d);

    if (stream) {
        stream_ = stream;
    } else {
        ETN_HIP(hipStreamCreate(&stream_));
        own_stream_ = true;
    }

    // BlockIterativeConfiguration: volume, mask, geometric setup, operators
    etn_volume_geometry vg = d.volume;
    if (vg.subvolume_count == 0) vg.subvolume_count = 1;   // the whole volume is resident in HBM
    volume_ = std::make_unique<GpuVolume>(vg, d.voxel_type);
    volume_->fill(d.init_value, stream_);
    mask_.data = d.mask;
    setup_ = std::make_unique<GeometricSetup>(geometries_[projection_set_->getProjectionIndex(0)],
                                              &volume_->properties());
    forward_ = std::make_unique<ForwardProjectionOperator>(setup_.get(), volume_.get(), &mask_, fp_samples_, stream_);
    backward_ = std::make_unique<BackProjectionOperator>(setup_.get(), volume_.get(), &mask_, lambda_, bp_samples_,
                                                         stream_);
    compare_ = std::make_unique<CompareKernel>(stream_);
    if (use_loc_) loc_ = std::make_unique<VirtualProjectionKernel>(stream_);

    // BlockIterativeState: 3 images per block slot
    const uint32_t slots = std::min(general_block_size_, projection_set_->getNumberOfProjections());
    state_.allocate(sizeof(float) * pixels_ * 3 * slots);
    for (uint32_t s = 0; s < slots; ++s) {
        float* base = state_.as<float>() + pixels_ * 3 * s;
        const uint32_t nu = d.base_geometry.resolution[0], nv = d.base_geometry.resolution[1];
        virtual_.push_back({base, nu, nv});
        traversal_.push_back({base + pixels_, nu, nv});
        residual_.push_back({base + 2 * pixels_, nu, nv});
    }
    // the norm runs over every measured projection not in the skip list
    for (uint32_t i = 0; i < d.projection_count; ++i)
        if (!skip.count(i)) measured_.push_back(i);
    for (uint32_t i = 0; i < d.projection_count; ++i)
        // (a 16-bit projection's address; the compare kernel reads it in projectionType())
        real_.push_back({static_cast<float*>(const_cast<void*>(realProjection(i))), d.base_geometry.resolution[0],
                         d.base_geometry.resolution[1]});
    norm_scratch_.allocate(sizeof(double) * (dev::kSsdBlocks + 1));

    if ((d.flags & ETN_RECON_FLAG_GENERIC_KERNELS) == 0) {
        std::vector<bool> taken;
        fused_ = FusedPlaneEngine::create({{this, &d}}, (d.flags & ETN_RECON_FLAG_EXACT_ARITHMETIC) != 0, stream_,
                                          &taken);
    }
    if ((d.flags & ETN_RECON_FLAG_STOCHASTIC_ROUNDING) && !fused_)
        fail("stochastic rounding needs the fast fused kernels: f16 or bf16 storage, no mask, no oversampling, "
             "single-axis tilt, nx a multiple of 4 and rays that cannot reach the 2048-step limit");
    tuned_ = (d.flags & ETN_RECON_FLAG_NO_AUTOTUNE) != 0;
}

BlockIterativeReconstructionOperator::~BlockIterativeReconstructionOperator()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    invalidate_graph();
    fused_.reset();
    if (own_stream_) (void)hipStreamDestroy(stream_);
}

void BlockIterativeReconstructionOperator::invalidate_graph()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
}

uint32_t BlockIterativeReconstructionOperator::getBlockSize(uint32_t block) const
{
    if (block + 1 < block_count_) return general_block_size_;
    if (block + 1 == block_count_)   // the last block holds the remainder
        return projection_set_->getNumberOfProjections() - (block_count_ - 1) * general_block_size_;
    return 0;
}

uint32_t BlockIterativeReconstructionOperator::getPhysicalProjectionIndex(uint32_t block, uint32_t p) const
{
    if (block >= block_count_ || p >= getBlockSize(block)) fail("Invalid projection identifier given!");
    return projection_set_->getProjectionIndex(general_block_size_ * block + p);
}

void BlockIterativeReconstructionOperator::run() { runNiterations(iterations_); }

void BlockIterativeReconstructionOperator::runSingleIteration() { runNiterations(1); }

void BlockIterativeReconstructionOperator::ensureTuned()
{
    if (tuned_ || !fused_) return;
    invalidate_graph();
    fused_->autotune(stream_);
    tuned_ = true;
}

void BlockIterativeReconstructionOperator::setTile(int planes, int band)
{
    if (!fused_) unsupported("this reconstruction does not use the fused plane kernels");
    invalidate_graph();
    fused_->setTile(0, planes, band, stream_);
    tuned_ = true;
}

void BlockIterativeReconstructionOperator::tile(int* planes, int* band) const
{
    if (!fused_) unsupported("this reconstruction does not use the fused plane kernels");
    fused_->tile(0, planes, band);
}

void BlockIterativeReconstructionOperator::setWorkgroupThreads(int threads)
{
    if (!fused_) unsupported("this reconstruction does not use the fused plane kernels");
    invalidate_graph();
    fused_->setThreads(threads, stream_);
    tuned_ = true;
}

int BlockIterativeReconstructionOperator::workgroupThreads() const
{
    if (!fused_) unsupported("this reconstruction does not use the fused plane kernels");
    return fused_->threads();
}

void BlockIterativeReconstructionOperator::runNiterations(uint32_t iterations)
{
    if (iterations == 0) return;
    ensureTuned();
    if (!use_graph_) {
        for (uint32_t i = 0; i < iterations; ++i) enqueueIteration();
        return;
    }
    if (!graph_exec_) {
        // Record one iteration and replay it: every iteration issues the same launches.
        ETN_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            enqueueIteration();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        ETN_HIP(hipStreamEndCapture(stream_, &graph_));
        ETN_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
    }
    for (uint32_t i = 0; i < iterations; ++i) ETN_HIP(hipGraphLaunch(graph_exec_, stream_));
}

void BlockIterativeReconstructionOperator::enqueueIteration()
{
    if (fused_) fused_->enqueueIteration(stream_);
    else handleOneIteration();
}

// BlockIterativeReconstructionOperator::handleOneIteration (:92-103)
void BlockIterativeReconstructionOperator::handleOneIteration()
{
    for (uint32_t block = 0; block < block_count_; ++block) {
        current_block_ = block;
        // BlockIterativeState::setCurrentBlockIndex clears the block's images (:29-48)
        for (uint32_t p = 0; p < getBlockSize(block); ++p)
            dev::launch_zero(virtual_[p].data, 3 * pixels_, stream_);
        doForwardProjectionForBlock(block);
        computeResidualsForBlock(block);
        doBackProjectionForBlock(block);
        if (on_block_end_) on_block_end_(block);
    }
}

void BlockIterativeReconstructionOperator::setBlockEndCallback(std::function<void(uint32_t block)> callback)
{
    if (callback && (fused_ || use_graph_))
        fail("a block-end callback needs the generic kernels launched eagerly");
    on_block_end_ = std::move(callback);
}

void BlockIterativeReconstructionOperator::doForwardProjectionForBlock(uint32_t block)
{
    const uint32_t subvolumes = volume_->properties().subvolume_count();
    const uint32_t bs = getBlockSize(block);
    for (uint32_t s = 0; s < subvolumes; ++s)
        for (uint32_t p = 0; p < bs; ++p) {
            current_projection_ = p;
            setup_->setScannerGeometry(geometries_[getPhysicalProjectionIndex(block, p)]);
            forward_->SetOutput(&virtual_[p], &traversal_[p]);
            forward_->Execute(s);
        }
}

void BlockIterativeReconstructionOperator::computeResidualsForBlock(uint32_t block)
{
    const uint32_t bs = getBlockSize(block);
    for (uint32_t p = 0; p < bs; ++p) {
        current_projection_ = p;
        if (use_loc_) {
            // the geometry last set by the FP: the block's last projection (:153-166)
            const float c = setup_->getScannerGeometry().traversal_length_through(volume_->properties().bbox_min(),
                                                                                  volume_->properties().bbox_max());
            loc_->setConstantLength(c);
            loc_->setProjection(&virtual_[p]);
            loc_->setTraversalLength(&traversal_[p]);
            loc_->run();
        }
        compare_->setInput(&real_[getPhysicalProjectionIndex(block, p)], &virtual_[p], projection_type_);
        compare_->setOutput(&residual_[p]);
        compare_->run();
    }
}

void BlockIterativeReconstructionOperator::doBackProjectionForBlock(uint32_t block)
{
    const uint32_t subvolumes = volume_->properties().subvolume_count();
    const uint32_t bs = getBlockSize(block);
    for (uint32_t s = 0; s < subvolumes; ++s)
        for (uint32_t p = 0; p < bs; ++p) {
            current_projection_ = p;
            setup_->setScannerGeometry(geometries_[getPhysicalProjectionIndex(block, p)]);
            backward_->setInput(&residual_[p], &traversal_[p], bs);
            backward_->execute(subvolumes - s - 1);
            backward_->setInput(nullptr, nullptr, 1);
        }
}

double BlockIterativeReconstructionOperator::residualNorm()
{
    if (fused_) {
        // the fused kernels' forward-only passes (one volume sweep per group
        // of projections instead of one generic launch per projection)
        const double norm = fused_->residualNorm(0, norm_scratch_.as<double>(), stream_);
        fused_->check(stream_);
        return norm;
    }
    double* partials = norm_scratch_.as<double>();
    double* total = partials + dev::kSsdBlocks;
    ETN_HIP(hipMemsetAsync(total, 0, sizeof(double), stream_));
    const uint32_t subvolumes = volume_->properties().subvolume_count();
    for (uint32_t phys : measured_) {
        dev::launch_zero(virtual_[0].data, 2 * pixels_, stream_);
        setup_->setScannerGeometry(geometries_[phys]);
        forward_->SetOutput(&virtual_[0], &traversal_[0]);
        for (uint32_t s = 0; s < subvolumes; ++s) forward_->Execute(s);
        dev::launch_sum_squared_difference(realProjection(phys), projection_type_, virtual_[0].data, pixels_, partials,
                                           total, stream_);
    }
    double host = 0.0;
    ETN_HIP(hipMemcpyAsync(&host, total, sizeof(double), hipMemcpyDeviceToHost, stream_));
    ETN_HIP(hipStreamSynchronize(stream_));
    return std::sqrt(host);
}

void BlockIterativeReconstructionOperator::synchronize()
{
    ETN_HIP(hipStreamSynchronize(stream_));
    if (fused_) fused_->check(stream_);
}

double BlockIterativeReconstructionOperator::timeKernels(int* kernel, uint32_t* launches)
{
    if (fused_) {
        // average over the launches that carry both a back and a forward projection
        ensureTuned();
        *kernel = 0;
        double total = 0.0;
        uint32_t n = 0, all = 0;
        for (const auto& t : fused_->timeIteration(stream_)) {
            ++all;
            if (t.steady) {
                total += t.ms;
                ++n;
            }
        }
        fused_->check(stream_);
        *launches = all;
        return n ? total / n : 0.0;
    }
    // one eager iteration with an event pair around every FP and BP launch
    std::vector<hipEvent_t> fp_ev, bp_ev;
    auto timed = [&](std::vector<hipEvent_t>& evs, auto&& launch) {
        hipEvent_t a, b;
        ETN_HIP(hipEventCreate(&a));
        ETN_HIP(hipEventCreate(&b));
        ETN_HIP(hipEventRecord(a, stream_));
        launch();
        ETN_HIP(hipEventRecord(b, stream_));
        evs.push_back(a);
        evs.push_back(b);
    };
    const uint32_t subvolumes = volume_->properties().subvolume_count();
    for (uint32_t block = 0; block < block_count_; ++block) {
        const uint32_t bs = getBlockSize(block);
        for (uint32_t p = 0; p < bs; ++p) dev::launch_zero(virtual_[p].data, 3 * pixels_, stream_);
        for (uint32_t s = 0; s < subvolumes; ++s)
            for (uint32_t p = 0; p < bs; ++p) {
                setup_->setScannerGeometry(geometries_[getPhysicalProjectionIndex(block, p)]);
                forward_->SetOutput(&virtual_[p], &traversal_[p]);
                timed(fp_ev, [&] { forward_->Execute(s); });
            }
        computeResidualsForBlock(block);
        for (uint32_t s = 0; s < subvolumes; ++s)
            for (uint32_t p = 0; p < bs; ++p) {
                setup_->setScannerGeometry(geometries_[getPhysicalProjectionIndex(block, p)]);
                backward_->setInput(&residual_[p], &traversal_[p], bs);
                timed(bp_ev, [&] { backward_->execute(subvolumes - s - 1); });
            }
    }
    ETN_HIP(hipStreamSynchronize(stream_));
    auto average = [](std::vector<hipEvent_t>& evs) {
        double total = 0.0;
        for (size_t i = 0; i < evs.size(); i += 2) {
            float ms = 0.0f;
            ETN_HIP(hipEventElapsedTime(&ms, evs[i], evs[i + 1]));
            total += ms;
        }
        for (auto e : evs) (void)hipEventDestroy(e);
        return evs.empty() ? 0.0 : total / (double)(evs.size() / 2);
    };
    const size_t n = fp_ev.size() / 2;
    const double fp = average(fp_ev), bp = average(bp_ev);
    *launches = (uint32_t)n;
    *kernel = fp >= bp ? 1 : 2;
    return fp >= bp ? fp : bp;
}

}  // namespace etn
