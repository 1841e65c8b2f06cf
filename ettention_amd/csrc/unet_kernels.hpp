// Kernels of the learned correction network (3-D U-Net): bf16 activations in
// [D][H][W][C] order, 3x3x3 convolutions as implicit GEMMs on the gfx950
// matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), and the
// elementwise / reduction kernels of training.  No reference counterpart.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace etn {
namespace nn {

// Input of a convolution: channels [0, ca) from `a` (at half resolution,
// read through nearest-neighbour upsampling, when a_half), then channels
// [ca, ca + cb) from `b` at full resolution.  ca and cb are multiples of 8
// (the network input is padded to 8 channels).
struct ConvSrc {
    const uint16_t* a = nullptr;
    int ca = 0;
    int a_half = 0;
    const uint16_t* b = nullptr;
    int cb = 0;
};

struct ConvArgs {
    ConvSrc src;
    int D = 0, H = 0, W = 0;      // output size (full resolution of the input)
    const uint16_t* wt = nullptr; // [cout][kpad] bf16, k = tap * cin + ci, zero past 27 * cin
    int kpad = 0;
    const float* bias = nullptr;  // [cout] or null
    int cout = 0;                 // 16, 32, 48, 64 or 96
    int relu = 0;
    uint16_t* out = nullptr;      // [D][H][W][cout] bf16
};

// Columns of the weight-gradient GEMM per workgroup.
constexpr int kWgradChunk = 512;

inline int kpad_for(int cin) { return (27 * cin + 31) / 32 * 32; }
inline int wgrad_chunks(size_t voxels) { return (int)((voxels + kWgradChunk - 1) / kWgradChunk); }

// y = act(conv3(src) + bias)
void conv3_forward(const ConvArgs& a, hipStream_t stream);
// partial[c][tap][co][ci16] = sum over voxels of chunk c of gy[m][co] * x[m + d(tap)][ci]
void conv3_weight_grad(const ConvSrc& src, int D, int H, int W, const uint16_t* gy, int cout, float* partial,
                       hipStream_t stream);
// dw[co][tap][ci] (fp32, scaled by `scale` and added to dw when accumulate) from the partials
void conv3_weight_grad_reduce(const float* partial, int chunks, int cout, int cin, float* dw, hipStream_t stream);
// partial[c][ch] = sum over the voxels of chunk c; then out[ch] = sum of partials
void channel_sum(const uint16_t* g, size_t voxels, int channels, float* partial, float* out, hipStream_t stream);

// packing of fp32 master weights [cout][27][cin] into the forward layout
// (wt[co][tap * cin_pad + ci], zero for ci >= cin) and the flipped, transposed
// backward-data layout (wb[ci][tap' * cout + co] = w[co][26 - tap'][ci]), both
// bf16 and zero padded to whole 32-wide k steps
void pack_weights(const float* w, int cout, int cin, int cin_pad, uint16_t* wt, uint16_t* wb, hipStream_t stream);

void relu_backward(uint16_t* g, const uint16_t* y, size_t count, hipStream_t stream);
void pool2(const uint16_t* x, int D, int H, int W, int C, uint16_t* y, hipStream_t stream);
// g_full[m][c] (+)= g_low[parent(m)][c] / 8
void pool2_backward(const uint16_t* g_low, int D, int H, int W, int C, uint16_t* g_full, int accumulate,
                    hipStream_t stream);
// g_low[p][c] = sum over the 8 children m of g[m][c0 + c], g has `stride` channels
void up2_backward(const uint16_t* g, int stride, int c0, int D, int H, int W, int C, uint16_t* g_low,
                  hipStream_t stream);
// g_dst[m][c] (+)= g[m][c0 + c]
void take_channels(const uint16_t* g, int stride, int c0, size_t voxels, int C, uint16_t* dst, int accumulate,
                   hipStream_t stream);

// y[m] = (x[m] scale, d[m] dscale, 0, 0, 0, 0, 0, 0) as eight bf16 channels; d may be null (zeros)
void pack_input(const float* x, float scale, const float* d, float dscale, size_t count, uint16_t* y,
                hipStream_t stream);
// pred[m] = x[m] * scale + (sum_c hw[c] d1[m][c] + hb), fp32
void head_forward(const uint16_t* d1, const float* x, float scale, const float* hw, const float* hb, size_t voxels,
                  float* pred, hipStream_t stream);
// out[m] = x[m] + (pred[m] - x[m] * scale) / scale
void unscale_prediction(const float* pred, const float* x, float scale, size_t voxels, float* out,
                        hipStream_t stream);
// r = pred - target * scale; partial sums of r^2 (fp64), g_pred = 2 r / voxels
void mse_loss(const float* pred, const float* target, float scale, size_t voxels, float* g_pred, double* partial,
              double* loss, hipStream_t stream);
// g_d1[m][c] = g_pred[m] hw[c]; partial head gradients (fp32) reduced into ghw[16], ghb[1]
void head_backward(const uint16_t* d1, const float* g_pred, const float* hw, size_t voxels, uint16_t* g_d1,
                   float* partial, float* ghw_ghb, hipStream_t stream);
void adam_step(float* p, const float* g, float* m, float* v, size_t count, float lr, float beta1, float beta2,
               float eps, float bc1, float bc2, hipStream_t stream);

}  // namespace nn
}  // namespace etn
