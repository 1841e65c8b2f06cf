// WBP (plugins/WBPPlugin/src/reconstruction/WBP.cpp:13-84): weighted back
// projection, the other caller of the back-projection operator.  Every
// projection is ramp-filtered (GenerateRamLakKernel + ConvolutionOperator),
// its pixel-centre chord through the volume box becomes the ray length
// (ComputeRayLengthKernel), and it is back-projected with lambda / N.
#pragma once

#include <memory>
#include <vector>

#include "fused.hpp"
#include "kernels.hpp"
#include "operators.hpp"
#include "reconstruction.hpp"

namespace etn {

class WeightedBackProjection final : public ReconstructionAlgorithm {
public:
    static constexpr uint32_t kFilterSize = 128;   // WBP.cpp:30

    WeightedBackProjection(const etn_reconstruction_desc& desc, hipStream_t stream);
    ~WeightedBackProjection() override;

    // WBP::run: every projection once, in stack order.
    void run() override;
    // WBP::runNiterations: projections 0 .. n-1.
    void runNiterations(uint32_t n) override;
    void runSingleIteration() override { runNiterations(1); }
    GpuVolume* getReconstructedVolume() override { return volume_.get(); }
    void synchronize() override;
    hipStream_t stream() const override { return stream_; }
    // Single-axis tilt geometries back-project on the fused kernels, eight
    // projections per volume sweep (fast arithmetic), unless the generic
    // kernels or exact arithmetic are asked for.
    bool usesFusedPath() const override { return fused_ != nullptr; }
    // One more run() with HIP events around the back projections: the average
    // fused launch carrying eight projections (*kernel = 0), or the generic
    // back projection (*kernel = 2).  Adds to the volume.
    double timeKernels(int* kernel, uint32_t* launches) override;

    // GenerateRamLakKernel.cl: value at position p = i - resolution / 2 is 1/4
    // at 0, 0 at even p, -1 / (p^2 pi^2) at odd p, in float.
    static std::vector<float> ramLakKernel(uint32_t resolution);
    // Filter parameters for an nu x nv image (ConvolutionFourierSpaceImplementation,
    // PaddingForFFT.cl, CroppingForFFT.cl).
    static dev::RampFilterParams filterParams(uint32_t nu, uint32_t nv, uint32_t taps);
    // ConvolutionOperator::execute on one device image.
    void filter(const float* src, float* dst);

private:
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    uint32_t count_;
    float lambda_;
    const float* projections_;
    size_t pixels_;
    std::vector<ParallelScannerGeometry> geometries_;
    std::unique_ptr<GpuVolume> volume_;
    GpuMask mask_;   // WBP passes no mask (WBP.cpp:45)
    std::unique_ptr<GeometricSetup> setup_;
    std::unique_ptr<BackProjectionOperator> backward_;
    DeviceBuffer images_;   // filtered projection + ray length
    DeviceImage filtered_, length_;
    std::unique_ptr<ComputeRayLengthKernel> ray_length_;
    DeviceBuffer taps_;
    dev::RampFilterParams params_;
    std::unique_ptr<FusedBackProjector> fused_;
    DeviceBuffer rays_;   // fused path: every projection's ray geometry, for the batched ray lengths
    std::vector<float>* launch_ms_ = nullptr;    // set while timeKernels runs
};

}  // namespace etn
