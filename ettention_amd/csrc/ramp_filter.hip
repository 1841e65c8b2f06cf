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

__global__ void __launch_bounds__(kRampThreads) ramp_filter_kernel(RampFilterParams p, const float* __restrict__ taps,
                                                                   const float* __restrict__ src,
                                                                   float* __restrict__ dst)
{
    __shared__ float row[kMaxRampRow];
    __shared__ float tap[kMaxRampTaps];
    const uint32_t v = blockIdx.x;
    // the shifted filter sits on padded row -1, so output row py reads padded row py + 1
    const int py = (int)((v + (uint32_t)p.crop_y + 1u) % p.uy);
    const int sy = mirror(py, p.pad_y, p.end_y, (int)p.nv);
    const float* s = src + (size_t)sy * p.nu;
    for (uint32_t x = threadIdx.x; x < p.ux; x += kRampThreads) row[x] = s[mirror((int)x, p.pad_x, p.end_x, (int)p.nu)];
    for (uint32_t t = threadIdx.x; t < p.taps; t += kRampThreads) tap[t] = taps[t];
    __syncthreads();
    const int half = (int)(p.taps / 2);
    const int ux = (int)p.ux;
    for (uint32_t u = threadIdx.x; u < p.nu; u += kRampThreads) {
        const int px = (int)u + p.crop_x;
        float acc = 0.0f;
        for (int t = 0; t < (int)p.taps; ++t) {
            int x = px - (t - half);   // |t - half| <= taps / 2 <= ux / 2: one wrap suffices
            x = x < 0 ? x + ux : (x >= ux ? x - ux : x);
            acc = fmaf(tap[t], row[x], acc);
        }
        dst[(size_t)v * p.nu + u] = acc;
    }
}

}  // namespace

void launch_ramp_filter(const RampFilterParams& p, const float* taps, const float* src, float* dst, hipStream_t stream)
{
    if (p.ux > kMaxRampRow || p.taps > kMaxRampTaps || p.taps > p.ux || p.nu > p.ux || p.nv > p.uy || p.nv == 0)
        fail("ramp filter: padded size out of range");
    ramp_filter_kernel<<<p.nv, kRampThreads, 0, stream>>>(p, taps, src, dst);
    ETN_HIP(hipGetLastError());
}

}  // namespace dev
}  // namespace etn
