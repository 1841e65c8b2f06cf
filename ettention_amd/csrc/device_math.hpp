// Device-side building blocks of the projectors: storage conversion, the DDA
// ray traversal and the software bilinear lookup.  Every file including this
// is compiled with -ffp-contract=off: the expressions below round exactly as
// the OpenCL C source of the reference writes them, so the traversal takes
// the same voxel sequence and the same segment lengths.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "kernels.hpp"

namespace etn {
namespace dev {

constexpr float kEpsilon = 1e-4f;     // ocl_src/Math.hl:4

// Global-memory pointers.  A pointer read from a descriptor table in memory is
// a generic pointer to the compiler, and accesses through it become flat
// instructions (issued to both the memory and the LDS pipes, results in
// VGPRs); marking it global gives global loads and scalar loads of uniform data.
// (The host pass of the single-source compile only parses device code: no
// address space there.)
// Marking a table constant (it is not written while the kernel runs) turns
// its uniform reads into scalar loads even after the kernel's own stores:
// a vector load of a per-projection parameter would otherwise wait (vmcnt, in
// issue order) for every band load issued before it.
#ifdef __HIP_DEVICE_COMPILE__
#define ETN_GLOBAL __attribute__((address_space(1)))
#define ETN_CONST __attribute__((address_space(4)))
#else
#define ETN_GLOBAL
#define ETN_CONST
#endif
template <typename T>
__device__ __forceinline__ ETN_GLOBAL T* global_ptr(T* p)
{
    return (ETN_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ const ETN_CONST T* const_ptr(const ETN_GLOBAL T* p)
{
    return (const ETN_CONST T*)(const T*)p;
}

// ---- voxel storage ---------------------------------------------------------
// load/store take any pointer to T (generic or global)
// Stochastic rounding (opt-in for 16-bit storage): f goes to one of the two
// representable values enclosing it, "down" (toward zero) and "up", and to
// "up" with probability (f - down) / (up - down), so the stored value is f in
// expectation: updates smaller than half a unit in the last place survive on
// average instead of being dropped.  r holds uniformly distributed bits.
__device__ __forceinline__ uint32_t sr_hash(uint32_t index, uint32_t salt, float value)
{
    // murmur3's 32-bit finaliser over the voxel index, the projection's salt
    // and the value being rounded (which changes from one update to the next)
    uint32_t h = index * 0x9E3779B1u ^ salt * 0x85EBCA77u ^ __float_as_uint(value);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

struct F32 {
    using T = float;
    template <typename P>
    static __device__ __forceinline__ float load(P p, size_t i) { return p[i]; }
    static __device__ __forceinline__ float round(float f) { return f; }
    static __device__ __forceinline__ float round_stochastic(float f, uint32_t) { return f; }
    template <typename P>
    static __device__ __forceinline__ void store(P p, size_t i, float f) { p[i] = f; }
};
struct F16 {   // vload_half / vstore_half (round to nearest even), HalfFloatSupport.hl:6-9
    using T = __half;
    template <typename P>
    static __device__ __forceinline__ float load(P p, size_t i)
    {
        const __half h = p[i];
        return __half2float(h);
    }
    static __device__ __forceinline__ float round(float f) { return __half2float(__float2half_rn(f)); }
    static __device__ __forceinline__ float round_stochastic(float f, uint32_t r)
    {
        // down: f rounded toward zero (exact when f is representable); up: the
        // next half away from zero (the sign kept; past the largest finite
        // half that is infinity, as round to nearest gives there too)
        const __half dn = __float2half_rz(f);
        const float d = __half2float(dn);
        const __half up = __ushort_as_half((unsigned short)(__half_as_ushort(dn) + 1u));
        const float u = __half2float(up);
        const float t = (float)(r >> 8) * 0x1.0p-24f;   // uniform in [0, 1)
        return fabsf(f - d) > t * fabsf(u - d) ? u : d;
    }
    template <typename P>
    static __device__ __forceinline__ void store(P p, size_t i, float f) { p[i] = __float2half_rn(f); }
    // two values in one v_cvt_pk_f16_f32 (gfx950, round to nearest even like
    // __float2half_rn): a in the low half, b in the high half
    typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ uint32_t pack2(float a, float b)
    {
        const f32x2 v = {a, b};
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2));
    }
    // the two halves of a packed pair, low half first (exact)
    static __device__ __forceinline__ void unpack2(uint32_t h, float& a, float& b)
    {
        const f32x2 v = __builtin_convertvector(__builtin_bit_cast(f16x2, h), f32x2);
        a = v.x;
        b = v.y;
    }
};
struct BF16 {  // bfloat16, round to nearest even
    using T = uint16_t;
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    // two values in one v_cvt_pk_bf16_f32 (gfx950's conversion, round to
    // nearest even; a NaN stays a NaN): a in the low half, b in the high half
    static __device__ __forceinline__ uint32_t pack2(float a, float b)
    {
        const f32x2 v = {a, b};
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    }
    static __device__ __forceinline__ uint16_t pack(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
    static __device__ __forceinline__ float unpack(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
    static __device__ __forceinline__ void unpack2(uint32_t h, float& a, float& b)
    {
        a = __uint_as_float(h << 16);
        b = __uint_as_float(h & 0xffff0000u);
    }
    template <typename P>
    static __device__ __forceinline__ float load(P p, size_t i) { return unpack(p[i]); }
    static __device__ __forceinline__ float round(float f) { return unpack(pack(f)); }
    static __device__ __forceinline__ float round_stochastic(float f, uint32_t r)
    {
        // add 16 random bits below the kept mantissa and truncate: the
        // magnitude rounds up with probability equal to the dropped fraction
        const uint32_t u = __float_as_uint(f);
        if ((u & 0x7fffffffu) >= 0x7f800000u) return f;   // infinity, NaN
        return __uint_as_float((u + (r & 0xffffu)) & 0xffff0000u);
    }
    template <typename P>
    static __device__ __forceinline__ void store(P p, size_t i, float f) { p[i] = pack(f); }
};

// ---- DDA ray traversal (ocl_src/DDARaytracer.hl:20-110, RayUtils.hl:7-35) ---

// Per-projection constants of the traversal.
struct DdaConst {
    float ray[3], inv[3], step[3];
    int sign[3];
};

__device__ __forceinline__ DdaConst dda_const(const float ray[3], const float vs[3])
{
    DdaConst c;
    for (int k = 0; k < 3; ++k) {
        c.ray[k] = ray[k];
        // invertVectorSafely (RayUtils.hl:25-35)
        c.inv[k] = 1.0f / ray[k];
        if (ray[k] == 0.0f) c.inv[k] = -FLT_MAX;
        const float sg = ray[k] > 0.0f ? 1.0f : (ray[k] < 0.0f ? -1.0f : 0.0f);
        c.step[k] = vs[k] * c.inv[k] * sg;    // stepDifference = voxelSize * invRay * sign(ray)
        c.sign[k] = (int)sg;
    }
    return c;
}

// Traversal state of one ray; suspends at band boundaries and resumes.
struct DdaRay {
    int vox[3];
    float res[3];
    float acc, accl;
    uint32_t brk;
    int live;
};

// Entry into the subvolume box; leaves the ray dead when it misses
// (getBoundingBoxEntryAndExitPointInRayCoordinates, doesRayIntersectBoundingBox,
// calculateRayEntryPoint, getInitialVoxelIndex, getRemainingDifferenceToVoxelBorder).
__device__ __forceinline__ void dda_begin(const SlabParams& s, const DdaConst& c, const float src[3], DdaRay& r)
{
    float tmin[3], tmax[3];
    for (int k = 0; k < 3; ++k) {
        const float t0 = (s.bmin[k] - src[k]) * c.inv[k];
        const float t1 = (s.bmax[k] - src[k]) * c.inv[k];
        tmin[k] = fminf(t0, t1);
        tmax[k] = fmaxf(t0, t1);
    }
    const float t_in = fmaxf(fmaxf(tmin[0], tmin[1]), tmin[2]);
    const float t_out = fminf(fminf(tmax[0], tmax[1]), tmax[2]);
    r.acc = 0.0f;
    r.accl = 0.0f;
    r.brk = 0;
    r.live = (t_out - t_in > kEpsilon) ? 1 : 0;
    if (!r.live) {
        r.vox[0] = r.vox[1] = r.vox[2] = -1;
        r.res[0] = r.res[1] = r.res[2] = 0.0f;
        return;
    }
    for (int k = 0; k < 3; ++k) {
        const float ivs = 1.0f / s.vs[k];
        float p = src[k] + t_in * c.ray[k];
        p = fminf(fmaxf(p, s.bmin[k]), s.bmax[k]);
        int vi = (int)((p - s.bmin[k]) * ivs);      // convert_int4_rtz
        vi = max(vi, 0);
        vi = min(vi, s.res[k] - 1);
        r.vox[k] = vi;
        float d = floorf(p * ivs) * s.vs[k];
        d = (c.ray[k] > 0.0f) ? (d + s.vs[k] - p) * c.inv[k] : (d - p) * c.inv[k];
        d = (c.ray[k] == 0.0f) ? FLT_MAX : d;
        d = (d == 0.0f) ? c.step[k] : d;
        r.res[k] = d;
    }
}

// Advances the ray while its voxel lies inside the subvolume and its z index
// inside [zlo, zhi).  `sample(x, y, z, &value, &maskv)` reads the voxel.  The
// loop condition is the reference's (in range, then ++breakVar < 4096, with a
// second ++breakVar in the body: at most 2048 steps); a ray that leaves the
// band with its condition not yet evaluated resumes unchanged.
template <typename Sampler>
__device__ __forceinline__ void dda_advance(const SlabParams& s, const DdaConst& c, DdaRay& r, int zlo, int zhi,
                                            const Sampler& sample)
{
    while (r.live) {
        if (!(r.vox[0] >= 0 && r.vox[1] >= 0 && r.vox[2] >= 0 && r.vox[0] < s.res[0] && r.vox[1] < s.res[1] &&
              r.vox[2] < s.res[2])) {
            r.live = 0;
            break;
        }
        if (r.vox[2] < zlo || r.vox[2] >= zhi) break;
        if (!(++r.brk < 4096u)) {
            r.live = 0;
            break;
        }
        const float dt = fminf(fminf(r.res[0], r.res[1]), r.res[2]);
        float value, maskv;
        sample(r.vox[0], r.vox[1], r.vox[2], value, maskv);
        r.acc += dt * value;
        r.accl += dt * maskv;
        const int ix = (r.res[0] <= r.res[1] && r.res[0] <= r.res[2]) ? 1 : 0;
        const int iy = (r.res[1] < r.res[0] && r.res[1] <= r.res[2]) ? 1 : 0;
        const int iz = (r.res[2] < r.res[0] && r.res[2] < r.res[1]) ? 1 : 0;
        r.res[0] = ix ? c.step[0] : r.res[0] - dt;
        r.res[1] = iy ? c.step[1] : r.res[1] - dt;
        r.res[2] = iz ? c.step[2] : r.res[2] - dt;
        r.vox[0] += ix * c.sign[0];
        r.vox[1] += iy * c.sign[1];
        r.vox[2] += iz * c.sign[2];
        ++r.brk;
    }
}

// Sub-ray origin (ParallelBeamsForwardProjection.cl:12-24).
__device__ __forceinline__ void fp_ray_source(const RayParams& g, uint32_t x, uint32_t y, uint32_t si, uint32_t sj,
                                              float sd, float off, float src[3])
{
    const float fx = (float)x + off + sd * (float)si;
    const float fy = (float)y + off + sd * (float)sj;
    for (int k = 0; k < 3; ++k) {
        const float ru = g.h[k] * fx;
        const float rv = g.v[k] * fy;
        src[k] = g.source[k] + ru + rv;
    }
}

// ---- bilinear lookup with zero border (ocl_src/ImageHandler.hl:20-40) ------
__device__ __forceinline__ float cl_fract(float x, float* base)
{
    const float fl = floorf(x);
    *base = fl;
    return fminf(x - fl, 0x1.fffffep-1f);
}
__device__ __forceinline__ float lerp_row(const float* row, int n, float c)
{
    float b;
    const float t = cl_fract(c - 0.5f, &b);
    const int i0 = (int)b, i1 = i0 + 1;
    const float v0 = (i0 >= 0 && i0 < n) ? row[i0] : 0.0f;
    const float v1 = (i1 >= 0 && i1 < n) ? row[i1] : 0.0f;
    return v0 + (v1 - v0) * t;
}
__device__ __forceinline__ float bilinear(const float* img, int nu, int nv, float cx, float cy)
{
    float b;
    const float t = cl_fract(cy - 0.5f, &b);
    const int j0 = (int)b, j1 = j0 + 1;
    const float v0 = (j0 >= 0 && j0 < nv) ? lerp_row(img + (size_t)j0 * nu, nu, cx) : 0.0f;
    const float v1 = (j1 >= 0 && j1 < nv) ? lerp_row(img + (size_t)j1 * nu, nu, cx) : 0.0f;
    return v0 + (v1 - v0) * t;
}

// Detector coordinate of a point (TransformCoordG, Matrix4x4.hl:38-58).
__device__ __forceinline__ void project_point(const float* m, const float p[3], float& u, float& v)
{
    float r[4];
    for (int row = 0; row < 4; ++row) {
        float a = m[0 * 4 + row] * p[0] + m[1 * 4 + row] * p[1];
        a = a + m[2 * 4 + row] * p[2];
        r[row] = a + m[3 * 4 + row] * 1.0f;
    }
    u = r[0] / r[3];
    v = r[1] / r[3];
}

// One voxel's correction: the sample loop of ParallelBeamsBackProjection.cl:26-44.
// Returns the used-sample count; *sum receives the summed contributions.
__device__ __forceinline__ uint32_t bp_voxel_sum(const BpParams& b, const float center[3], const float* residual,
                                                 const float* ray_length, float* sum)
{
    float acc = 0.0f;
    uint32_t used = 0;
    for (uint32_t si = 0; si < b.nsamples; ++si) {
        const float4 smp = b.samples[si];
        const float p[3] = {center[0] + smp.x, center[1] + smp.y, center[2] + smp.z};
        float u, v;
        project_point(b.m, p, u, v);
        const float L = bilinear(ray_length, (int)b.nu, (int)b.nv, u, v);
        const float R = bilinear(residual, (int)b.nu, (int)b.nv, u, v);
        const float w = smp.w / L;
        const float contrib = R * b.lambda * w;
        acc += (L > kEpsilon) ? contrib : 0.0f;
        used += (L > kEpsilon) ? 1u : 0u;
    }
    *sum = acc;
    return used;
}

}  // namespace dev
}  // namespace etn
