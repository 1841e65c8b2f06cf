#include "batch.hpp"
#include "kernels.hpp"

namespace etn {

BatchReconstruction::BatchReconstruction(const etn_reconstruction_desc* members, uint32_t count, uint32_t flags,
                                         hipStream_t stream)
    : use_graph_((flags & ETN_RECON_FLAG_NO_GRAPH) == 0), tuned_((flags & ETN_RECON_FLAG_NO_AUTOTUNE) != 0)
{
    if (count == 0 || !members) fail("a batch needs at least one member");
    if (stream) {
        stream_ = stream;
    } else {
        ETN_HIP(hipStreamCreate(&stream_));
        own_stream_ = true;
    }
    try {
        // every member on the batch stream, with its own kernels disabled: the
        // batch decides which members the fused engines take
        std::vector<etn_reconstruction_desc> descs(members, members + count);
        for (auto& d : descs) {
            const etn_reconstruction_desc own = d;
            d.flags = (own.flags & ~(uint32_t)ETN_RECON_FLAG_STOCHASTIC_ROUNDING) | ETN_RECON_FLAG_GENERIC_KERNELS |
                      ETN_RECON_FLAG_NO_GRAPH | ETN_RECON_FLAG_NO_AUTOTUNE;
            ops_.push_back(std::make_unique<BlockIterativeReconstructionOperator>(d, stream_));
        }
        slots_.assign(count, Slot{});
        if ((flags & ETN_RECON_FLAG_GENERIC_KERNELS) == 0) {
            // one engine per (voxel type, row-vector width) among the eligible members
            std::vector<uint32_t> left;
            for (uint32_t i = 0; i < count; ++i)
                if ((members[i].flags & ETN_RECON_FLAG_GENERIC_KERNELS) == 0) left.push_back(i);
            while (!left.empty()) {
                std::vector<FusedPlaneEngine::Member> cand;
                for (uint32_t i : left) cand.push_back({ops_[i].get(), &members[i]});
                std::vector<bool> taken;
                std::unique_ptr<FusedPlaneEngine> e =
                    FusedPlaneEngine::create(cand, (flags & ETN_RECON_FLAG_EXACT_ARITHMETIC) != 0, stream_, &taken);
                if (!e) break;
                std::vector<uint32_t> rest;
                size_t index = 0;
                for (size_t k = 0; k < left.size(); ++k) {
                    if (taken[k]) slots_[left[k]] = {(int)engines_.size(), index++};
                    else rest.push_back(left[k]);
                }
                engines_.push_back(std::move(e));
                left.swap(rest);
            }
        }
        for (uint32_t i = 0; i < count; ++i)
            if ((members[i].flags & ETN_RECON_FLAG_STOCHASTIC_ROUNDING) && slots_[i].engine < 0)
                fail("batch member " + std::to_string(i) + ": stochastic rounding needs the fast fused kernels "
                     "(f16 or bf16 storage, no mask, no oversampling, nx a multiple of 4, rays that cannot "
                     "reach the 2048-step limit)");
    } catch (...) {
        ops_.clear();
        engines_.clear();
        if (own_stream_) (void)hipStreamDestroy(stream_);
        throw;
    }
}

BatchReconstruction::~BatchReconstruction()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    invalidateGraph();
    engines_.clear();
    ops_.clear();
    if (own_stream_) (void)hipStreamDestroy(stream_);
}

void BatchReconstruction::invalidateGraph()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
}

BlockIterativeReconstructionOperator& BatchReconstruction::member(uint32_t i)
{
    if (i >= ops_.size()) fail("batch member index out of range");
    return *ops_[i];
}

bool BatchReconstruction::usesFusedPath(uint32_t i) const
{
    if (i >= ops_.size()) fail("batch member index out of range");
    return slots_[i].engine >= 0;
}

void BatchReconstruction::tile(uint32_t i, int* planes, int* band) const
{
    if (!usesFusedPath(i)) unsupported("this batch member does not use the fused plane kernels");
    engines_[slots_[i].engine]->tile(slots_[i].index, planes, band);
}

double BatchReconstruction::residualNorm(uint32_t i)
{
    if (i >= ops_.size()) fail("batch member index out of range");
    if (slots_[i].engine < 0) return ops_[i]->residualNorm();
    if (!norm_scratch_.get()) norm_scratch_.allocate(sizeof(double) * (dev::kSsdBlocks + 1));
    FusedPlaneEngine& e = *engines_[slots_[i].engine];
    const double norm = e.residualNorm(slots_[i].index, norm_scratch_.as<double>(), stream_);
    e.check(stream_);
    return norm;
}

int BatchReconstruction::workgroupThreads(uint32_t i) const
{
    if (!usesFusedPath(i)) unsupported("this batch member does not use the fused plane kernels");
    return engines_[slots_[i].engine]->threads();
}

void BatchReconstruction::ensureTuned()
{
    if (tuned_) return;
    invalidateGraph();
    for (auto& e : engines_) e->autotune(stream_);
    tuned_ = true;
}

void BatchReconstruction::enqueueIteration()
{
    for (auto& e : engines_) e->enqueueIteration(stream_);
    for (uint32_t i = 0; i < ops_.size(); ++i)
        if (slots_[i].engine < 0) ops_[i]->enqueueIteration();
}

void BatchReconstruction::runNiterations(uint32_t iterations)
{
    if (iterations == 0) return;
    ensureTuned();
    if (!use_graph_) {
        for (uint32_t i = 0; i < iterations; ++i) enqueueIteration();
        return;
    }
    if (!graph_exec_) {
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

void BatchReconstruction::synchronize()
{
    ETN_HIP(hipStreamSynchronize(stream_));
    for (auto& e : engines_) e->check(stream_);
}

double BatchReconstruction::timeIteration(uint32_t* launches)
{
    ensureTuned();
    double total = 0.0;
    uint32_t n = 0;
    for (auto& e : engines_)
        for (const auto& t : e->timeIteration(stream_)) {
            total += t.ms;
            ++n;
        }
    for (uint32_t i = 0; i < ops_.size(); ++i)
        if (slots_[i].engine < 0) ops_[i]->enqueueIteration();
    synchronize();
    *launches = n;
    return total;
}

}  // namespace etn
