// Ramp (RamLak) filtering of a projection for weighted back projection.
// The reference convolves in Fourier space with clFFT; the same circular
// convolution is summed directly here, one workgroup per detector row, the
// padded source row staged in LDS.
#include "common.hpp"
#include "kernels.hpp"

namespace etn {
namespace dev {

namespace {

constexpr int kRampThreads = 256;

// PaddingForFFT.cl:computePaddedCopyForFourierSpace maps a padded coordinate
// to the image: inside as is, before the image mirrored about its first
// sample, after it mirrored about one past its last.  The reference reads
// outside the image when the padding is wider than the image; here the index
// is clamped.
__device__ __forceinline__ int mirror(int x, int begin, int end, int size)
{
    int m;
    if (x >= begin && x < end) m = x - begin;
    else if (x < end) m = begin - x;
    else m = (size + end - 1) - x;
    return min(max(m, 0), size - 1);
}

// One workgroup per detector row (blockIdx.x) of image blockIdx.y.
__global__ void __launch_bounds__(kRampThreads) ramp_filter_kernel(RampFilterParams p, const float* __restrict__ taps,
                                                                   const float* __restrict__ src, size_t src_stride,
                                                                   float* __restrict__ dst, size_t dst_stride)
{
    __shared__ float row[kMaxRampRow];
    __shared__ float tap[kMaxRampTaps];
    const uint32_t v = blockIdx.x;
    src += (size_t)blockIdx.y * src_stride;
    dst += (size_t)blockIdx.y * dst_stride;
    // the shifted filter sits on padded row -1, so output row py reads padded row py + 1
    const int py = (int)((v + (uint32_t)p.crop_y + 1u) % p.uy);
    const int sy = mirror(py, p.pad_y, p.end_y, (int)p.nv);
    const float* s = src + (size_t)sy * p.nu;
    for (uint32_t x = threadIdx.x; x < p.ux; x += kRampThreads) row[x] = s[mirror((int)x, p.pad_x, p.end_x, (int)p.nu)];
    for (uint32_t t = threadIdx.x; t < p.taps; t += kRampThreads) tap[t] = taps[t];
    __syncthreads();
    const int half = (int)(p.taps / 2);
    const int ux = (int)p.ux;
    // |t - half| <= taps / 2 <= ux / 2 and the cropped columns lie in the
    // padded row: one wrap suffices
    auto at = [&](int x) { return row[x < 0 ? x + ux : (x >= ux ? x - ux : x)]; };
    // four consecutive outputs per thread: tap t reads padded column
    // px - (t - half) for each, so the four columns slide down by one per tap
    // and one new column is read per tap (the sums run over t in order, as one
    // output per thread would)
    const int groups = (int)((p.nu + 3) / 4);
    for (int g = (int)threadIdx.x; g < groups; g += kRampThreads) {
        const int px0 = 4 * g + p.crop_x;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = at(px0 + j + half);
        for (int t = 0; t < (int)p.taps; ++t) {
            const float c = tap[t];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j] = fmaf(c, w[j], acc[j]);
            w[3] = w[2];
            w[2] = w[1];
            w[1] = w[0];
            w[0] = at(px0 - (t + 1) + half);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (4 * g + j < (int)p.nu) dst[(size_t)v * p.nu + 4 * g + j] = acc[j];
    }
}

}  // namespace

void launch_ramp_filter_batch(const RampFilterParams& p, const float* taps, const float* src, size_t src_stride,
                              float* dst, size_t dst_stride, uint32_t count, hipStream_t stream)
{
    if (p.ux > kMaxRampRow || p.taps > kMaxRampTaps || p.taps > p.ux || p.nu > p.ux || p.nv > p.uy || p.nv == 0)
        fail("ramp filter: padded size out of range");
    if (count == 0) return;
    if (count > 65535) fail("ramp filter: at most 65535 images per launch");
    ramp_filter_kernel<<<dim3(p.nv, count), kRampThreads, 0, stream>>>(p, taps, src, src_stride, dst, dst_stride);
    ETN_HIP(hipGetLastError());
}

void launch_ramp_filter(const RampFilterParams& p, const float* taps, const float* src, float* dst, hipStream_t stream)
{
    launch_ramp_filter_batch(p, taps, src, 0, dst, 0, 1, stream);
}

}  // namespace dev
}  // namespace etn
