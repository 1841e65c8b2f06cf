// Learned correction network: a 3-level 3-D U-Net (16/32/64 channels, 3x3x3
// convolutions with ReLU, 2x2x2 average pooling, nearest upsampling with
// channel-concatenated skips, a 1x1x1 head) that maps an early SART iterate x
// and its last update d (x minus the iterate before it) to a late iterate as
// x + delta(x, d).  Weights and activations are bf16 on the
// matrix cores with FP32 accumulation; master weights, gradients and the Adam
// state are FP32.  Parameter order (flat vector): per layer c1a c1b c2a c2b
// c3a c3b c4a c4b c5a c5b the weights [cout][27][cin] then the bias [cout];
// then the head weights [16] and bias.  No reference counterpart.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"
#include "unet_kernels.hpp"

namespace etn {

class UNet3D {
public:
    struct Layer {
        const char* name;
        int cin, cout;
        int level;            // 0 full, 1 half, 2 quarter resolution
        size_t w, b;          // offsets into the parameter vector
    };

    // Working size nx x ny x nz (multiples of 4); volumes are [nz][ny][nx].
    UNet3D(int nx, int ny, int nz, hipStream_t stream);

    static size_t parameterCount();
    void setParameters(const float* host, size_t count);
    void getParameters(float* host, size_t count) const;

    // out = x + delta(x scale, d dscale) / scale (device fp32 volumes; d may be null: zeros)
    void predict(const float* x, const float* d, float scale, float dscale, float* out);
    // forward + backward of L = mean((x s + delta - target s)^2); gradients stay on the device
    double gradient(const float* x, const float* d, const float* target, float scale, float dscale);
    void copyGradient(float* host, size_t count) const;
    // gradient + one Adam step (beta1 0.9, beta2 0.999, eps 1e-8)
    double trainStep(const float* x, const float* d, const float* target, float scale, float dscale, float lr);
    // HIP-event time of one forward pass; flops = 2 x multiply-adds of its convolutions
    double timeForward(const float* x, float scale, double* flops);

private:
    void forward(const float* x, const float* d, float scale, float dscale);
    void backward();
    void packAll();
    void ensureTraining();
    nn::ConvSrc layerInput(int i) const;
    const uint16_t* layerOutput(int i) const { return act_[i].as<uint16_t>(); }
    size_t voxels(int level) const { return (size_t)D_[level] * H_[level] * W_[level]; }

    hipStream_t stream_;
    int D_[3], H_[3], W_[3];
    std::vector<Layer> layers_;
    size_t head_;                                   // offset of the head weights
    DeviceBuffer params_, grads_, adam_m_, adam_v_;
    std::vector<DeviceBuffer> wt_, wb_;             // packed bf16 weights per layer
    DeviceBuffer xin_, pool1_, pool2_, pred_;
    std::vector<DeviceBuffer> act_;                 // output of each layer (bf16)
    // training buffers
    bool training_ = false;
    DeviceBuffer gpred_, gd1_, gtmp_, gcat_, ge1_, ge2_, ge3_, gd2_, gp_, wpart_, bpart_, hpart_, loss_, lpart_;
    int adam_t_ = 0;
};

}  // namespace etn
