// BlockIterativeReconstructionOperator: the SART / SIRT / block-iterative loop
// (ettention/libReconstruction/src/algorithm/reconstruction/blockiterative/).
#pragma once

#include <functional>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "operators.hpp"
#include "projection_set.hpp"
#include "reconstruction.hpp"

namespace etn {

class FusedPlaneEngine;

class BlockIterativeReconstructionOperator final : public ReconstructionAlgorithm {
public:
    BlockIterativeReconstructionOperator(const etn_reconstruction_desc& desc, hipStream_t stream);
    ~BlockIterativeReconstructionOperator() override;

    // ReconstructionAlgorithm (ReconstructionAlgorithm.h:9-22)
    void run() override;
    void runSingleIteration() override;
    void runNiterations(uint32_t iterations) override;
    GpuVolume* getReconstructedVolume() override { return volume_.get(); }
    // The fused plane kernels take the mask given at creation; with them a
    // different mask is refused (create the reconstruction with it instead).
    void setMaskVolume(const uint8_t* mask)
    {
        if (fused_ && mask != mask_.data) unsupported("the fused plane kernels keep the mask given at creation");
        mask_.data = mask;
        invalidate_graph();
    }
    const uint8_t* getMaskVolume() const { return mask_.data; }

    // Extension: ||A x - b||_2 over the used projections, FP64 accumulation (synchronises).
    double residualNorm() override;
    void synchronize() override;
    // One eager iteration with HIP events around every launch; returns the
    // average milliseconds of the dominant kernel (0 fused, 1 FP, 2 BP).
    double timeKernels(int* kernel, uint32_t* launches) override;
    hipStream_t stream() const override { return stream_; }
    bool usesFusedPath() const override { return fused_ != nullptr; }
    void invalidate_graph();
    // Fused path only: runs the tile autotuner once (unless disabled or a tile was set).
    void ensureTuned() override;
    void setTile(int planes, int band) override;
    void tile(int* planes, int* band) const override;
    void setWorkgroupThreads(int threads) override;
    int workgroupThreads() const override;
    // The launches of one iteration on stream(), without graph capture.
    void enqueueIteration();
    // Called on the host after each block's back projection has been enqueued
    // (the logger's OnBlockEnd hook, BlockIterativeReconstructionAlgorithmLogger.cpp:99-110).
    // Needs the generic kernels launched eagerly (ETN_RECON_FLAG_GENERIC_KERNELS |
    // ETN_RECON_FLAG_NO_GRAPH): the fused kernels overlap consecutive blocks.
    void setBlockEndCallback(std::function<void(uint32_t block)> callback);

    // BlockIterativeProjectionAccess (BlockIterativeProjectionAccess.cpp:11-110)
    uint32_t GetBlockCount() const { return block_count_; }
    uint32_t getBlockSize(uint32_t block) const;
    uint32_t getPhysicalProjectionIndex(uint32_t block, uint32_t p) const;
    const ParallelScannerGeometry& geometry(uint32_t physical) const { return geometries_.at(physical); }
    // measured projection `physical` in its storage type (projectionType())
    const void* realProjection(uint32_t physical) const
    {
        return static_cast<const char*>(projections_) + (size_t)physical * pixels_ * voxel_bytes(projection_type_);
    }
    etn_voxel_type projectionType() const { return projection_type_; }
    // Physical indices of the projections the residual norm runs over (all but the skip list).
    const std::vector<uint32_t>& measuredProjections() const { return measured_; }

private:
    void handleOneIteration();
    void doForwardProjectionForBlock(uint32_t block);
    void computeResidualsForBlock(uint32_t block);
    void doBackProjectionForBlock(uint32_t block);

    etn_reconstruction_desc desc_;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    uint32_t iterations_;
    float lambda_;
    bool use_loc_;
    uint32_t fp_samples_, bp_samples_;
    uint32_t general_block_size_;
    uint32_t block_count_ = 0;
    std::unique_ptr<ProjectionSet> projection_set_;
    std::vector<ParallelScannerGeometry> geometries_;     // per physical projection
    const void* projections_;
    etn_voxel_type projection_type_ = ETN_VOXEL_FLOAT32;   // f16/bf16 with ETN_RECON_FLAG_16BIT_PROJECTIONS
    size_t pixels_;

    std::unique_ptr<GpuVolume> volume_;
    GpuMask mask_;
    std::unique_ptr<GeometricSetup> setup_;
    std::unique_ptr<ForwardProjectionOperator> forward_;
    std::unique_ptr<BackProjectionOperator> backward_;
    std::unique_ptr<CompareKernel> compare_;
    std::unique_ptr<VirtualProjectionKernel> loc_;

    // BlockIterativeState: per-slot virtual projection, traversal length, residual
    DeviceBuffer state_;
    std::vector<DeviceImage> virtual_, traversal_, residual_;
    std::vector<DeviceImage> real_;
    std::vector<uint32_t> measured_;
    uint32_t current_block_ = 0, current_projection_ = 0;

    DeviceBuffer norm_scratch_;
    std::unique_ptr<FusedPlaneEngine> fused_;

    bool use_graph_;
    bool tuned_ = true;
    std::function<void(uint32_t)> on_block_end_;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
};

}  // namespace etn
