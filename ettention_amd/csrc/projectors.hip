// Generic gfx950 kernels of the SART hot path: one thread per detector pixel
// (forward projection, residual, long-object compensation, ray length) or per
// voxel (back projection), for any parallel geometry, any subvolume split and
// any sample count.  The fused plane kernels (fused.hip) take over the common
// single-axis case; these serve everything else and the per-kernel C ABI.
#include <mutex>
#include <vector>

#include "device_math.hpp"
#include "common.hpp"

namespace etn {
namespace dev {

namespace {

template <typename S, bool kMask>
struct GlobalSampler {
    const typename S::T* vol;
    const uint8_t* mask;
    uint64_t base;
    uint32_t nx, nxy;
    __device__ __forceinline__ void operator()(int x, int y, int z, float& value, float& maskv) const
    {
        const size_t i = base + (size_t)z * nxy + (size_t)y * nx + (size_t)x;
        value = S::load(vol, i);
        maskv = kMask ? (float)mask[i] / 255.0f : 1.0f;
    }
};

// ParallelBeamsForwardProjection.cl:execute
template <typename S, bool kMask>
__global__ void __launch_bounds__(256) forward_project_kernel(SlabParams s, RayParams g, const typename S::T* vol,
                                                              const uint8_t* mask, float* proj, float* len)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.nu || y >= g.nv) return;
    const DdaConst c = dda_const(g.ray, s.vs);
    const GlobalSampler<S, kMask> sample{vol, mask, s.first_voxel, s.nx, s.nx * s.ny};
    const uint32_t n = g.samples;
    const float total = (float)(n * n);
    const float sd = 1.0f / (float)n, off = sd / 2.0f;
    float acc = 0.0f, accl = 0.0f;
    for (uint32_t si = 0; si < n; ++si)
        for (uint32_t sj = 0; sj < n; ++sj) {
            float src[3];
            fp_ray_source(g, x, y, si, sj, sd, off, src);
            DdaRay r;
            dda_begin(s, c, src, r);
            dda_advance(s, c, r, 0, s.res[2], sample);
            acc += r.acc;
            accl += r.accl;
        }
    const size_t i = (size_t)y * g.nu + x;
    proj[i] += acc / total;
    len[i] += accl / total;
}

// ParallelBeamsBackProjection.cl:correct
template <typename S, bool kMask>
__global__ void __launch_bounds__(256) back_project_kernel(SlabParams s, BpParams b, typename S::T* vol,
                                                           const uint8_t* mask, const float* residual,
                                                           const float* ray_length)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= s.res[0] || y >= s.res[1] || z >= s.res[2]) return;
    const float center[3] = {s.bmin[0] + ((float)x + 0.5f) * s.vs[0], s.bmin[1] + ((float)y + 0.5f) * s.vs[1],
                             s.bmin[2] + ((float)z + 0.5f) * s.vs[2]};
    float sum;
    const uint32_t used = bp_voxel_sum(b, center, residual, ray_length, &sum);
    const size_t i = s.first_voxel + ((size_t)z * s.ny + (size_t)y) * s.nx + (size_t)x;
    float value = S::load(vol, i);
    if (used != 0) {
        const float maskv = kMask ? (float)mask[i] / 255.0f : 1.0f;
        value += maskv * sum / (float)used;
    }
    S::store(vol, i, value);
}

// the measured projection in its storage type (f32, or f16/bf16 stored projections)
template <typename S>
__global__ void compare_kernel(uint32_t n, const typename S::T* real, const float* virt, float* res)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) res[i] = S::load(real, i) - virt[i];
}

__global__ void loc_kernel(uint32_t n, float* proj, float* len, float c)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float l = len[i];
    proj[i] = l > kEpsilon ? proj[i] * c / l : 0.0f;
    len[i] = c;
}

// ComputeRayLength.cl:execute for pixel (x, y)
__device__ __forceinline__ float ray_length_at(const RayParams& g, const float bmin[3], const float bmax[3],
                                               uint32_t x, uint32_t y)
{
    float src[3];
    const float fx = ((float)x) + 0.5f, fy = ((float)y) + 0.5f;
    float tmin[3], tmax[3];
    for (int k = 0; k < 3; ++k) {
        const float ru = g.h[k] * fx, rv = g.v[k] * fy;
        src[k] = g.source[k] + ru + rv;
        float inv = 1.0f / g.ray[k];
        if (g.ray[k] == 0.0f) inv = -FLT_MAX;
        const float t0 = (bmin[k] - src[k]) * inv, t1 = (bmax[k] - src[k]) * inv;
        tmin[k] = fminf(t0, t1);
        tmax[k] = fmaxf(t0, t1);
    }
    const float t_in = fmaxf(fmaxf(tmin[0], tmin[1]), tmin[2]);
    const float t_out = fminf(fminf(tmax[0], tmax[1]), tmax[2]);
    return fmaxf(t_out - t_in, 0.0f);
}

__global__ void ray_length_kernel(RayParams g, float bx0, float by0, float bz0, float bx1, float by1, float bz1,
                                  float* out)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.nu || y >= g.nv) return;
    const float bmin[3] = {bx0, by0, bz0}, bmax[3] = {bx1, by1, bz1};
    out[(size_t)y * g.nu + x] = ray_length_at(g, bmin, bmax, x, y);
}

// Image blockIdx.z of a batch: its geometry from gs, its lengths at out + z * stride.
__global__ void ray_length_batch_kernel(const RayParams* __restrict__ gs, float bx0, float by0, float bz0, float bx1,
                                        float by1, float bz1, float* out, size_t stride)
{
    const RayParams g = gs[blockIdx.z];
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.nu || y >= g.nv) return;
    const float bmin[3] = {bx0, by0, bz0}, bmax[3] = {bx1, by1, bz1};
    out[(size_t)blockIdx.z * stride + (size_t)y * g.nu + x] = ray_length_at(g, bmin, bmax, x, y);
}

// FP64 sum of squared differences: per-block partials, then one block sums
// the partials in a fixed order (deterministic).
constexpr int kSsdThreads = 256;
template <typename S>
__global__ void __launch_bounds__(kSsdThreads) ssd_partial_kernel(const typename S::T* a, const float* b, size_t n,
                                                                  double* partials)
{
    __shared__ double red[kSsdThreads];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double d = (double)S::load(a, i) - (b ? (double)b[i] : 0.0);
        acc += d * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kSsdThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(kSsdThreads) ssd_final_kernel(const double* partials, uint32_t n, double* out)
{
    __shared__ double red[kSsdThreads];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kSsdThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += red[0];
}

template <typename SS, typename DS>
__global__ void convert_kernel(const typename SS::T* src, typename DS::T* dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        DS::store(dst, i, SS::load(src, i));
}

template <typename S>
__global__ void fill_kernel(typename S::T* dst, float value, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        S::store(dst, i, value);
}

uint32_t grid_for(size_t n, uint32_t threads)
{
    const size_t blocks = (n + threads - 1) / threads;
    return (uint32_t)(blocks < 8192 ? (blocks ? blocks : 1) : 8192);
}

void check_launch(const char* what)
{
    hip_check(hipGetLastError(), what);
}

}  // namespace

void launch_forward_project(const SlabParams& s, const RayParams& r, etn_voxel_type vt, const void* volume,
                            const uint8_t* mask, float* projection, float* traversal_length, hipStream_t stream)
{
    const dim3 block(64, 4);
    const dim3 grid((r.nu + 63) / 64, (r.nv + 3) / 4);
    switch (vt) {
        case ETN_VOXEL_FLOAT32:
            if (mask)
                forward_project_kernel<F32, true><<<grid, block, 0, stream>>>(s, r, (const float*)volume, mask, projection, traversal_length);
            else
                forward_project_kernel<F32, false><<<grid, block, 0, stream>>>(s, r, (const float*)volume, mask, projection, traversal_length);
            break;
        case ETN_VOXEL_FLOAT16:
            if (mask)
                forward_project_kernel<F16, true><<<grid, block, 0, stream>>>(s, r, (const __half*)volume, mask, projection, traversal_length);
            else
                forward_project_kernel<F16, false><<<grid, block, 0, stream>>>(s, r, (const __half*)volume, mask, projection, traversal_length);
            break;
        case ETN_VOXEL_BFLOAT16:
            if (mask)
                forward_project_kernel<BF16, true><<<grid, block, 0, stream>>>(s, r, (const uint16_t*)volume, mask, projection, traversal_length);
            else
                forward_project_kernel<BF16, false><<<grid, block, 0, stream>>>(s, r, (const uint16_t*)volume, mask, projection, traversal_length);
            break;
    }
    check_launch("forward projection launch");
}

void launch_back_project(const SlabParams& s, const BpParams& b, etn_voxel_type vt, void* volume,
                         const uint8_t* mask, const float* residual, const float* ray_length, hipStream_t stream)
{
    const dim3 block(64, 4);
    const dim3 grid((s.res[0] + 63) / 64, (s.res[1] + 3) / 4, s.res[2]);
    switch (vt) {
        case ETN_VOXEL_FLOAT32:
            if (mask)
                back_project_kernel<F32, true><<<grid, block, 0, stream>>>(s, b, (float*)volume, mask, residual, ray_length);
            else
                back_project_kernel<F32, false><<<grid, block, 0, stream>>>(s, b, (float*)volume, mask, residual, ray_length);
            break;
        case ETN_VOXEL_FLOAT16:
            if (mask)
                back_project_kernel<F16, true><<<grid, block, 0, stream>>>(s, b, (__half*)volume, mask, residual, ray_length);
            else
                back_project_kernel<F16, false><<<grid, block, 0, stream>>>(s, b, (__half*)volume, mask, residual, ray_length);
            break;
        case ETN_VOXEL_BFLOAT16:
            if (mask)
                back_project_kernel<BF16, true><<<grid, block, 0, stream>>>(s, b, (uint16_t*)volume, mask, residual, ray_length);
            else
                back_project_kernel<BF16, false><<<grid, block, 0, stream>>>(s, b, (uint16_t*)volume, mask, residual, ray_length);
            break;
    }
    check_launch("back projection launch");
}

void launch_compare(uint32_t n, const void* real, etn_voxel_type real_type, const float* virt, float* residual,
                    hipStream_t stream)
{
    const uint32_t grid = (n + 255) / 256;
    if (real_type == ETN_VOXEL_FLOAT16)
        compare_kernel<F16><<<grid, 256, 0, stream>>>(n, (const __half*)real, virt, residual);
    else if (real_type == ETN_VOXEL_BFLOAT16)
        compare_kernel<BF16><<<grid, 256, 0, stream>>>(n, (const uint16_t*)real, virt, residual);
    else
        compare_kernel<F32><<<grid, 256, 0, stream>>>(n, (const float*)real, virt, residual);
    check_launch("compare launch");
}

void launch_compare(uint32_t n, const float* real, const float* virt, float* residual, hipStream_t stream)
{
    launch_compare(n, real, ETN_VOXEL_FLOAT32, virt, residual, stream);
}

void launch_long_object_compensate(uint32_t n, float* projection, float* length, float c, hipStream_t stream)
{
    loc_kernel<<<(n + 255) / 256, 256, 0, stream>>>(n, projection, length, c);
    check_launch("long object compensation launch");
}

void launch_ray_length(const float bmin[3], const float bmax[3], const RayParams& r, float* out, hipStream_t stream)
{
    const dim3 block(64, 4);
    const dim3 grid((r.nu + 63) / 64, (r.nv + 3) / 4);
    ray_length_kernel<<<grid, block, 0, stream>>>(r, bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2], out);
    check_launch("ray length launch");
}

void launch_ray_length_batch(const float bmin[3], const float bmax[3], const RayParams* params, uint32_t count,
                             uint32_t nu, uint32_t nv, float* out, size_t stride, hipStream_t stream)
{
    if (count == 0) return;
    if (count > 65535) fail("ray length: at most 65535 images per launch");
    const dim3 block(64, 4);
    const dim3 grid((nu + 63) / 64, (nv + 3) / 4, count);
    ray_length_batch_kernel<<<grid, block, 0, stream>>>(params, bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2],
                                                        out, stride);
    check_launch("ray length launch");
}

void launch_sum_squared_difference(const void* a, etn_voxel_type a_type, const float* b, size_t count,
                                   double* partials, double* out, hipStream_t stream)
{
    if (a_type == ETN_VOXEL_FLOAT16)
        ssd_partial_kernel<F16><<<kSsdBlocks, kSsdThreads, 0, stream>>>((const __half*)a, b, count, partials);
    else if (a_type == ETN_VOXEL_BFLOAT16)
        ssd_partial_kernel<BF16><<<kSsdBlocks, kSsdThreads, 0, stream>>>((const uint16_t*)a, b, count, partials);
    else
        ssd_partial_kernel<F32><<<kSsdBlocks, kSsdThreads, 0, stream>>>((const float*)a, b, count, partials);
    ssd_final_kernel<<<1, kSsdThreads, 0, stream>>>(partials, kSsdBlocks, out);
    check_launch("residual norm launch");
}

// Scratch of the allocation-free residual norm: the kSsdBlocks FP64 partial
// sums of ssd_partial_kernel (4 KiB), one copy per GPU, held by each GPU the
// library runs on until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host it names a stand-in, not device memory): it
// looks up each device's address once, with that device current, and uses it.
static __device__ __align__(256) double scratch[kSsdBlocks];

double* ssd_static_scratch(int device)
{
    static std::mutex lock;
    static std::vector<double*> address;   // per device, filled on first use
    std::lock_guard<std::mutex> guard(lock);
    if (device < 0) fail("residual norm scratch: invalid device");
    if ((size_t)device >= address.size()) address.resize((size_t)device + 1, nullptr);
    if (!address[device]) {
        int current = 0;
        ETN_HIP(hipGetDevice(&current));
        if (current != device) ETN_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device) ETN_HIP(hipSetDevice(current));
        ETN_HIP(e);
        if (!p) fail("residual norm scratch: no device address");
        address[device] = static_cast<double*>(p);
    }
    return address[device];
}

void launch_sum_squared_difference(const float* a, const float* b, size_t count, double* partials, double* out,
                                   hipStream_t stream)
{
    launch_sum_squared_difference(a, ETN_VOXEL_FLOAT32, b, count, partials, out, stream);
}

void launch_convert(const void* src, etn_voxel_type st, void* dst, etn_voxel_type dt, size_t count, hipStream_t stream)
{
    const uint32_t grid = grid_for(count, 256);
#define ETN_CONVERT(SS, DS)                                                                                    \
    convert_kernel<SS, DS><<<grid, 256, 0, stream>>>((const typename SS::T*)src, (typename DS::T*)dst, count)
    if (st == ETN_VOXEL_FLOAT32 && dt == ETN_VOXEL_FLOAT32) ETN_CONVERT(F32, F32);
    else if (st == ETN_VOXEL_FLOAT32 && dt == ETN_VOXEL_FLOAT16) ETN_CONVERT(F32, F16);
    else if (st == ETN_VOXEL_FLOAT32 && dt == ETN_VOXEL_BFLOAT16) ETN_CONVERT(F32, BF16);
    else if (st == ETN_VOXEL_FLOAT16 && dt == ETN_VOXEL_FLOAT32) ETN_CONVERT(F16, F32);
    else if (st == ETN_VOXEL_FLOAT16 && dt == ETN_VOXEL_FLOAT16) ETN_CONVERT(F16, F16);
    else if (st == ETN_VOXEL_FLOAT16 && dt == ETN_VOXEL_BFLOAT16) ETN_CONVERT(F16, BF16);
    else if (st == ETN_VOXEL_BFLOAT16 && dt == ETN_VOXEL_FLOAT32) ETN_CONVERT(BF16, F32);
    else if (st == ETN_VOXEL_BFLOAT16 && dt == ETN_VOXEL_FLOAT16) ETN_CONVERT(BF16, F16);
    else ETN_CONVERT(BF16, BF16);
#undef ETN_CONVERT
    check_launch("convert launch");
}

void launch_fill(void* dst, etn_voxel_type t, float value, size_t count, hipStream_t stream)
{
    const uint32_t grid = grid_for(count, 256);
    if (t == ETN_VOXEL_FLOAT32) fill_kernel<F32><<<grid, 256, 0, stream>>>((float*)dst, value, count);
    else if (t == ETN_VOXEL_FLOAT16) fill_kernel<F16><<<grid, 256, 0, stream>>>((__half*)dst, value, count);
    else fill_kernel<BF16><<<grid, 256, 0, stream>>>((uint16_t*)dst, value, count);
    check_launch("fill launch");
}

void launch_zero(float* dst, size_t count, hipStream_t stream)
{
    fill_kernel<F32><<<grid_for(count, 256), 256, 0, stream>>>(dst, 0.0f, count);
    check_launch("zero launch");
}

}  // namespace dev
}  // namespace etn
