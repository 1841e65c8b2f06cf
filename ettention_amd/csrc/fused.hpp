// Fused plane engine: runs whole block-iterative iterations with kernels that
// keep whole volume planes (fixed y) per workgroup and fuse the back projection
// of block b-1 with the forward projection of block b in one pass over the
// volume.  Used when every ray stays in a y plane (single-axis tilt about y).
//
// One engine drives any number of independent reconstructions ("members") of
// possibly different sizes: each launch carries one pass of every member that
// still has passes left, as a table of (member, plane group) workgroups.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"

namespace etn {

class BlockIterativeReconstructionOperator;
class GpuVolume;
class ParallelScannerGeometry;

// Back projections of many projections into one volume with the fused
// kernels' back-projection phase alone (fast arithmetic): each launch applies
// up to eight projections in one sweep of the volume, in list order, rounding
// to the storage type after each as separate launches would.  For callers of
// the BP other than the block-iterative loop (weighted back projection).
class FusedBackProjector {
public:
    // Null when the geometries have no fused form (x-axis tilt, nx not a
    // multiple of 4, a sample count the fused kernels do not take).
    // lds_bp: apply the projections to the band in LDS one after another,
    // never in registers as the band is loaded (same volumes).
    static std::unique_ptr<FusedBackProjector> create(GpuVolume* volume,
                                                      const std::vector<ParallelScannerGeometry>& geometries,
                                                      uint32_t nu, uint32_t nv, uint32_t bp_samples, bool lds_bp,
                                                      hipStream_t stream);
    virtual ~FusedBackProjector() = default;
    // Back-projects projections indices[k] (into the geometry list) with
    // residual[k] and ray length[k] (device nv x nu images) and weight
    // lambda (already divided by the block size).  Synchronises.  With
    // `launch_ms`, appends the event-timed duration of every launch that
    // carries eight projections.
    virtual void backProject(const std::vector<uint32_t>& indices, const std::vector<const float*>& residual,
                             const std::vector<const float*>& length, float lambda, hipStream_t stream,
                             std::vector<float>* launch_ms = nullptr) = 0;
};

class FusedPlaneEngine {
public:
    struct Member {
        BlockIterativeReconstructionOperator* op;
        const etn_reconstruction_desc* desc;
    };
    struct LaunchTime {
        float ms;
        bool steady;            // every member pass of the launch carries a back and a forward projection
    };

    // Builds an engine over the members that have a fused form (geometry,
    // samples, subvolumes and mask inside what the fused kernels cover) and
    // share the first such member's voxel type and row-vector width.
    // taken[i] tells whether members[i] is driven by the engine; null when
    // none is.  exact: the generic kernels' float operations in their order
    // (bit-identical volumes); otherwise the reference's operations with
    // combined bilinear weights, a hardware reciprocal and fused multiply-adds.
    static std::unique_ptr<FusedPlaneEngine> create(const std::vector<Member>& members, bool exact,
                                                    hipStream_t stream, std::vector<bool>* taken);
    virtual ~FusedPlaneEngine() = default;
    virtual void enqueueIteration(hipStream_t stream) = 0;
    // Throws when a kernel found a live ray outside the plane its row was assigned to.
    virtual void check(hipStream_t stream) = 0;
    // One eager iteration with an event pair around every launch.
    virtual std::vector<LaunchTime> timeIteration(hipStream_t stream) = 0;
    // Times the candidate tiles (planes per workgroup, z rows per band) of
    // every member size on the members' own data, for each workgroup shape,
    // and keeps the fastest.  The volumes are restored afterwards; tiles never
    // change results.
    virtual void autotune(hipStream_t stream) = 0;
    // Index is the member's position among the engine's members (in the order
    // given to create).  Throws ETN_ERROR_INVALID_ARGUMENT for a tile that
    // does not fit; reallocates the launch tables (no captured graph may use them).
    virtual void setTile(size_t member, int planes, int band, hipStream_t stream) = 0;
    virtual void tile(size_t member, int* planes, int* band) const = 0;
    // Workgroup shape of every launch: 512 threads (64 KiB LDS, two workgroups
    // per CU) or 256 (40 KiB, four per CU).  The autotuner picks one; results
    // never depend on it.  Tiles that no longer fit revert to the default.
    virtual void setThreads(int threads, hipStream_t stream) = 0;
    virtual int threads() const = 0;
    virtual size_t memberCount() const = 0;
    // ||A x - b||_2 of a member over its measured projections, FP64: the
    // forward projections run as forward-only passes of the fused kernels,
    // the squared differences are summed per projection in the member's
    // projection order.  `scratch` holds kSsdBlocks + 1 doubles.  Synchronises.
    virtual double residualNorm(size_t member, double* scratch, hipStream_t stream) = 0;
};

}  // namespace etn
