// Batched reconstructions: many independent block-iterative reconstructions
// of different sizes advanced together.  No reference counterpart: the
// reference runs one reconstruction per process.
#pragma once

#include <memory>
#include <vector>

#include "block_iterative.hpp"
#include "fused.hpp"

namespace etn {

class BatchReconstruction {
public:
    BatchReconstruction(const etn_reconstruction_desc* members, uint32_t count, uint32_t flags, hipStream_t stream);
    ~BatchReconstruction();

    void runNiterations(uint32_t iterations);
    void synchronize();
    uint32_t size() const { return (uint32_t)ops_.size(); }
    BlockIterativeReconstructionOperator& member(uint32_t i);
    bool usesFusedPath(uint32_t i) const;
    void tile(uint32_t i, int* planes, int* band) const;
    int workgroupThreads(uint32_t i) const;   // workgroup shape of the member's fused launches
    // ||A x - b||_2 of member i (FP64; on its engine's fused kernels when it has one).  Synchronises.
    double residualNorm(uint32_t i);
    // One eager iteration: total milliseconds of the fused launches and their count.
    double timeIteration(uint32_t* launches);

private:
    struct Slot {
        int engine = -1;        // index into engines_, -1: the member's own kernels
        size_t index = 0;       // member index inside its engine
    };
    void ensureTuned();
    void enqueueIteration();
    void invalidateGraph();

    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    bool use_graph_, tuned_;
    std::vector<std::unique_ptr<BlockIterativeReconstructionOperator>> ops_;
    std::vector<std::unique_ptr<FusedPlaneEngine>> engines_;
    std::vector<Slot> slots_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    DeviceBuffer norm_scratch_;     // FP64 partial sums of the residual norm
};

}  // namespace etn
