// ReconstructionAlgorithm (ettention/libReconstruction/src/algorithm/reconstruction/
// ReconstructionAlgorithm.h:9-22): what a caller of a reconstruction sees,
// whichever algorithm the parameters selected.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace etn {

class GpuVolume;

class ReconstructionAlgorithm {
public:
    virtual ~ReconstructionAlgorithm() = default;
    virtual void run() = 0;
    virtual void runSingleIteration() = 0;
    virtual void runNiterations(uint32_t iterations) = 0;
    virtual GpuVolume* getReconstructedVolume() = 0;
    virtual void synchronize() = 0;
    virtual hipStream_t stream() const = 0;

    // Extensions of the block-iterative loop; other algorithms refuse them.
    virtual double residualNorm() { unsupported("this reconstruction has no residual norm"); }
    virtual bool usesFusedPath() const { return false; }
    virtual double timeKernels(int*, uint32_t*) { unsupported("this reconstruction has no kernel timing"); }
    virtual void ensureTuned() {}
    virtual void setTile(int, int) { unsupported("this reconstruction does not use the fused plane kernels"); }
    virtual void tile(int*, int*) const { unsupported("this reconstruction does not use the fused plane kernels"); }
    virtual void setWorkgroupThreads(int) { unsupported("this reconstruction does not use the fused plane kernels"); }
    virtual int workgroupThreads() const { unsupported("this reconstruction does not use the fused plane kernels"); }
};

}  // namespace etn
