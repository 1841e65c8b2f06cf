// Fused plane engine: runs whole block-iterative iterations with kernels that
// keep one volume plane (fixed y) per workgroup and fuse the back projection
// of block b-1 with the forward projection of block b in one pass over the
// volume.  Used when every ray stays in a y plane (single-axis tilt about y).
#pragma once

#include <memory>

#include "common.hpp"

namespace etn {

class BlockIterativeReconstructionOperator;

class FusedPlaneEngine {
public:
    // Null when the reconstruction has no fused form (geometry, samples,
    // subvolumes or mask outside what the fused kernels cover).
    static std::unique_ptr<FusedPlaneEngine> create(BlockIterativeReconstructionOperator& op,
                                                    const etn_reconstruction_desc& desc, hipStream_t stream);
    virtual ~FusedPlaneEngine() = default;
    virtual void enqueueIteration(hipStream_t stream) = 0;
    // Throws when a kernel found a live ray outside the plane its row was assigned to.
    virtual void check(hipStream_t stream) = 0;
    // One eager iteration with events around every launch: average milliseconds of
    // the launches that carry both a back and a forward projection.
    virtual double timeKernels(hipStream_t stream, uint32_t* launches) = 0;
};

}  // namespace etn
