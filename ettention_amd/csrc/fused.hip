// Fused plane kernels for single-axis tilt geometries.
#include "fused.hpp"

#include "block_iterative.hpp"

namespace etn {

std::unique_ptr<FusedPlaneEngine> FusedPlaneEngine::create(BlockIterativeReconstructionOperator&,
                                                           const etn_reconstruction_desc&, hipStream_t)
{
    return nullptr;
}

}  // namespace etn
