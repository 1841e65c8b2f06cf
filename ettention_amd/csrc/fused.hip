// Fused plane kernels for single-axis tilt geometries (rays with no y
// component, detector rows of constant y).
//
// Every ray then stays in one volume plane y = j, and the back projection of
// a voxel in plane j reads detector rows that are constant across the plane.
// One workgroup owns one plane.  It streams the plane through LDS in bands of
// z rows, in the direction the rays travel in z, and for each band
//   1. applies the back projections of the previous block (in block order,
//      rounding to the storage type after each, exactly as the per-projection
//      launches do) and writes the band back to HBM, then
//   2. advances the rays of the current block's forward projections through
//      the band (DDA state kept in registers across bands).
// After the last band each ray's line integral becomes the residual (and
// optional long-object compensation) of its pixel.  One pass therefore reads
// and writes the volume once where the unfused schedule reads it three times
// (FP, BP read, BP write); the arithmetic per ray and per voxel is the generic
// kernels', operation for operation, so results are bit-identical to them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "block_iterative.hpp"
#include "device_math.hpp"
#include "fused.hpp"

namespace etn {
namespace dev {

constexpr int kFusedThreads = 512;
constexpr int kMaxFusedFp = 16;      // forward projections per pass
constexpr int kMaxFusedBp = 8;       // back projections per pass
constexpr int kMaxRaysPerThread = 8;

struct FusedFp {
    float src[3], h[3], v[3], ray[3];
    const float* measured;       // nv x nu
    float* residual;             // nv x nu, written
    float* length;               // nv x nu, written
    float loc_c;
    int use_loc;
};

struct FusedBp {
    float m[16];
    float lambda;                // lambda / block size
    const float* residual;       // nv x nu
    const float* length;         // nv x nu
    float4 sample;               // single BP sample: offset xyz, in-voxel chord w
};

struct FusedPass {
    int nfp, nbp;
    int zdir;                    // +1: bands in increasing z, -1: decreasing
    int pad;
    FusedFp fp[kMaxFusedFp];
    FusedBp bp[kMaxFusedBp];
};

struct FusedVolume {
    void* data;
    SlabParams slab;             // the single subvolume (bounding box as VolumeProperties computes it)
    int nx, ny, nz;
    int nu, nv;
    int band;                    // z rows per LDS band
    const int* row_lo;           // per plane: first detector row whose rays lie in it
    const int* row_hi;           // per plane: one past the last such row
    const int* dead_rows;        // detector rows whose rays all miss the volume in y
    int ndead;
    int* diag;                   // counts live rays found outside their plane (must stay 0)
};

namespace {

// DDA state of a ray confined to plane y = j.  The y index never steps and
// res.y stays FLT_MAX, which turns the reference's step choice
//   ix = rx <= ry && rx <= rz,  iz = rz < rx && rz < ry,  dt = min(rx, ry, rz)
// into exactly ix = rx <= rz, iz = !ix, dt = min(rx, rz): rx and rz are never
// NaN and never above FLT_MAX (an axis with zero direction keeps FLT_MAX and is
// never chosen; every other value is a finite difference).
struct PlaneRay {
    int vx, vz;
    float rx, rz;
    float acc, accl;
    uint32_t brk;
    int live;
    float stepx, stepz;
    int sx, sz;
};

// Advances a ray while it is inside the volume and inside band [zlo, zhi).  A
// ray leaving the volume is finished; one leaving the band resumes in the next
// band with unchanged state.  kCap keeps the reference's 2048-step limit; it is
// dropped when nx + nz <= 2049, since every step moves vx or vz monotonically by
// one, so a ray can take at most nx + nz - 1 steps inside the volume.
template <bool kCap>
__device__ __forceinline__ void plane_advance(PlaneRay& r, const float* band, int nx, int nz, int zlo, int zhi)
{
    if (!r.live) return;
    const unsigned h = (unsigned)(zhi - zlo);
    const int lz = r.sz * nx;
    int li = (r.vz - zlo) * nx + r.vx;
    int vx = r.vx, vz = r.vz;
    float rx = r.rx, rz = r.rz, acc = r.acc, accl = r.accl;
    for (;;) {
        const bool inx = (unsigned)vx < (unsigned)nx;
        if (!(inx && (unsigned)(vz - zlo) < h)) {
            if (!inx || (unsigned)vz >= (unsigned)nz) r.live = 0;
            break;
        }
        if (kCap) {
            if (!(++r.brk < 4096u)) {
                r.live = 0;
                break;
            }
        }
        // dt = min(rx, rz) as a compare-select: it differs from fminf at most in
        // the sign of a zero, which changes neither the sums nor the next res values
        const bool ix = rx <= rz;
        const float dt = ix ? rx : rz;
        acc += dt * band[li];
        accl += dt;                                     // dt * 1.0f
        const float rest = (ix ? rz : rx) - dt;         // the axis not stepped loses dt
        rx = ix ? r.stepx : rest;
        rz = ix ? rest : r.stepz;
        vx += ix ? r.sx : 0;
        vz += ix ? 0 : r.sz;
        li += ix ? r.sx : lz;
        if (kCap) ++r.brk;
    }
    r.vx = vx;
    r.vz = vz;
    r.rx = rx;
    r.rz = rz;
    r.acc = acc;
    r.accl = accl;
}

// Four-voxel (or one-voxel) groups of a band row in the storage type: HBM
// access granularity of the band transfers.
template <typename S, bool kVec>
struct GroupIo;
template <>
struct GroupIo<F32, true> {
    static constexpr int G = 4;
    using V = float4;
    static __device__ __forceinline__ V load(const float* p, size_t i) { return *reinterpret_cast<const float4*>(p + i); }
    static __device__ __forceinline__ void unpack(const V& v, float* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
    static __device__ __forceinline__ void store(float* p, size_t i, const float* f)
    {
        *reinterpret_cast<float4*>(p + i) = make_float4(f[0], f[1], f[2], f[3]);
    }
};
template <typename S16>
struct GroupIo16 {   // 4 x 16-bit voxels as one 8-byte access
    static constexpr int G = 4;
    using V = uint2;
    static __device__ __forceinline__ V load(const typename S16::T* p, size_t i) { return *reinterpret_cast<const uint2*>(p + i); }
    static __device__ __forceinline__ void unpack(const V& v, float* o)
    {
        o[0] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 0);
        o[1] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 1);
        o[2] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 2);
        o[3] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 3);
    }
    static __device__ __forceinline__ void store(typename S16::T* p, size_t i, const float* f)
    {
        V v;
        typename S16::T* t = reinterpret_cast<typename S16::T*>(&v);
        for (int k = 0; k < 4; ++k) S16::store(t, k, f[k]);
        *reinterpret_cast<uint2*>(p + i) = v;
    }
};
template <>
struct GroupIo<F16, true> : GroupIo16<F16> {};
template <>
struct GroupIo<BF16, true> : GroupIo16<BF16> {};
template <typename S>
struct GroupIo<S, false> {
    static constexpr int G = 1;
    using V = typename S::T;
    static __device__ __forceinline__ V load(const typename S::T* p, size_t i) { return p[i]; }
    static __device__ __forceinline__ void unpack(const V& v, float* o) { o[0] = S::load(&v, 0); }
    static __device__ __forceinline__ void store(typename S::T* p, size_t i, const float* f) { S::store(p, i, f[0]); }
};

constexpr int kLdsFloats = 16384;   // 64 KiB per workgroup
constexpr int kRowPad = 2;          // zero taps on each side of a staged detector row

// LinearInterpolation (ImageHandler.hl:20-29) on a staged row with kRowPad zeros on
// each side, stored as (v[i], v[i+1] - v[i]) pairs: clamping the base index to
// [-2, n] reads only zeros wherever the reference's range checks return 0, and
// the stored difference is the subtraction mix() performs, so the value is the same.
__device__ __forceinline__ int padded_index(int i0, int n) { return min(max(i0, -kRowPad), n) + kRowPad; }
__device__ __forceinline__ float lerp_pair(float2 p, float t) { return p.x + p.y * t; }

template <typename S, bool kVec, bool kCap, int RPT>
__global__ void __launch_bounds__(kFusedThreads) fused_plane_kernel(FusedVolume fv, const FusedPass* __restrict__ pass)
{
    using IO = GroupIo<S, kVec>;
    constexpr int G = IO::G;
    constexpr int kGroups = kLdsFloats / (G * kFusedThreads);   // band groups per thread (upper bound)
    extern __shared__ float lds[];
    const int j = blockIdx.x;
    const int tid = threadIdx.x;
    const int nx = fv.nx, nz = fv.nz, nu = fv.nu;
    const int row_len = nu + 2 * kRowPad;
    float* band = lds;                                  // band x nx
    // 4 padded detector rows of (value, difference), 8-byte aligned after the band
    float2* rows = reinterpret_cast<float2*>(lds + ((fv.band * nx + 1) & ~1));
    typename S::T* vol = static_cast<typename S::T*>(fv.data);
    const size_t plane_stride = (size_t)nx * fv.ny;
    const size_t plane_base = (size_t)j * nx;
    const int gpr = nx / G;                             // groups per band row

    const int nfp = pass->nfp, nbp = pass->nbp;
    const int row_lo = fv.row_lo[j], nrows = fv.row_hi[j] - row_lo;
    const int rays_per_proj = nrows * nu;
    const int nrays = nfp * rays_per_proj;
    const SlabParams& s = fv.slab;

    // ---- ray set-up: dda_begin exactly as the generic FP kernel ----
    PlaneRay ray[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int id = k * kFusedThreads + tid;
        ray[k].live = 0;
        ray[k].acc = 0.0f;
        ray[k].accl = 0.0f;
        if (id < nrays) {
            const int q = id / rays_per_proj;
            const int rem = id - q * rays_per_proj;
            const int w = row_lo + rem / nu;
            const int u = rem - (rem / nu) * nu;
            const FusedFp& f = pass->fp[q];
            RayParams g;
            for (int a = 0; a < 3; ++a) {
                g.source[a] = f.src[a];
                g.h[a] = f.h[a];
                g.v[a] = f.v[a];
                g.ray[a] = f.ray[a];
            }
            const DdaConst c = dda_const(f.ray, s.vs);
            float src[3];
            fp_ray_source(g, (uint32_t)u, (uint32_t)w, 0, 0, 1.0f, 0.5f, src);
            DdaRay r;
            dda_begin(s, c, src, r);
            if (r.live && r.vox[1] != j) {
                atomicAdd(fv.diag, 1);
                r.live = 0;
            }
            ray[k].vx = r.vox[0];
            ray[k].vz = r.vox[2];
            ray[k].rx = r.res[0];
            ray[k].rz = r.res[2];
            ray[k].brk = r.brk;
            ray[k].live = r.live;
            ray[k].stepx = c.step[0];
            ray[k].stepz = c.step[2];
            ray[k].sx = c.sign[0];
            ray[k].sz = c.sign[2];
        }
    }

    const int nbands = (nz + fv.band - 1) / fv.band;
    const float py = s.bmin[1] + ((float)j + 0.5f) * s.vs[1];

    // detector rows of back projection q: the v coordinate is constant over the
    // plane (checked on the host), so rows j0, j0+1 and the weight ty are too
    int j0 = 0;
    float ty = 0.0f;
    auto stage_rows = [&](int q) {
        const FusedBp& bp = pass->bp[q];
        float vp = bp.m[1] * 0.0f + bp.m[5] * (py + bp.sample.y);
        vp = vp + bp.m[9] * 0.0f;
        vp = vp + bp.m[13] * 1.0f;
        float base_row;
        ty = cl_fract(vp - 0.5f, &base_row);
        j0 = (int)base_row;
        for (int e = tid; e < 2 * row_len; e += kFusedThreads) {
            const int rr = e / row_len, k = e - rr * row_len, u = k - kRowPad;
            const int row = j0 + rr;
            const bool rok = row >= 0 && row < fv.nv;
            const bool ok0 = rok && u >= 0 && u < nu, ok1 = rok && u + 1 >= 0 && u + 1 < nu;
            const float r0 = ok0 ? bp.residual[(size_t)row * nu + u] : 0.0f;
            const float r1 = ok1 ? bp.residual[(size_t)row * nu + u + 1] : 0.0f;
            const float l0 = ok0 ? bp.length[(size_t)row * nu + u] : 0.0f;
            const float l1 = ok1 ? bp.length[(size_t)row * nu + u + 1] : 0.0f;
            rows[rr * row_len + k] = make_float2(r0, r1 - r0);
            rows[(2 + rr) * row_len + k] = make_float2(l0, l1 - l0);
        }
    };

    auto band_range = [&](int bi, int& z0, int& z1) {
        const int b = pass->zdir >= 0 ? bi : nbands - 1 - bi;
        z0 = b * fv.band;
        z1 = min(z0 + fv.band, nz);
    };

    // the vector path keeps the next band in registers while this band is
    // processed; the scalar path (nx not a multiple of 4) loads at commit time
    constexpr int kPre = kVec ? kGroups : 1;
    typename IO::V pre[kPre];
    auto prefetch = [&](int z0, int z1) {
        if (!kVec) return;
        const int ng = (z1 - z0) * gpr;
#pragma unroll
        for (int k = 0; k < kPre; ++k) {
            const int gi = k * kFusedThreads + tid;
            if (gi < ng) {
                const int zz = gi / gpr, xg = gi - zz * gpr;
                pre[k] = IO::load(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xg * G);
            }
        }
    };

    int z0, z1;
    band_range(0, z0, z1);
    prefetch(z0, z1);
    if (nbp == 1) stage_rows(0);     // made visible by the first band's barrier

    for (int bi = 0; bi < nbands; ++bi) {
        band_range(bi, z0, z1);
        const int ng = (z1 - z0) * gpr;
        // 1. prefetched band -> LDS, then start fetching the next band
        if (kVec) {
#pragma unroll
            for (int k = 0; k < kPre; ++k) {
                const int gi = k * kFusedThreads + tid;
                if (gi < ng) {
                    float v[G];
                    IO::unpack(pre[k], v);
                    const int zz = gi / gpr, xg = gi - zz * gpr;
#pragma unroll
                    for (int i = 0; i < G; ++i) band[zz * nx + xg * G + i] = v[i];
                }
            }
        } else {
            for (int gi = tid; gi < ng; gi += kFusedThreads) {
                const int zz = gi / gpr, xg = gi - zz * gpr;
                float v[G];
                IO::unpack(IO::load(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xg * G), v);
#pragma unroll
                for (int i = 0; i < G; ++i) band[zz * nx + xg * G + i] = v[i];
            }
        }
        __syncthreads();
        if (bi + 1 < nbands) {
            int n0, n1;
            band_range(bi + 1, n0, n1);
            prefetch(n0, n1);
        }

        // 2. back projections of the previous block, in block order; each thread
        //    updates (and finally stores) the groups it staged
        for (int q = 0; q < nbp; ++q) {
            if (nbp > 1) {
                stage_rows(q);
                __syncthreads();
            }
            const FusedBp& bp = pass->bp[q];
            const float2* r0 = rows;
            const float2* r1 = rows + row_len;
            const float2* l0 = rows + 2 * row_len;
            const float2* l1 = rows + 3 * row_len;
            const float m0 = bp.m[0], m8 = bp.m[8], m12 = bp.m[12], lam = bp.lambda, chord = bp.sample.w;
            const float a4 = bp.m[4] * (py + bp.sample.y);
            // u = ((m0 px + m4 py) + m8 pz) + m12 w, w == 1 exactly (checked on the host)
            auto x_term = [&](int x) {
                const float px = s.bmin[0] + ((float)x + 0.5f) * s.vs[0] + bp.sample.x;
                return m0 * px + a4;
            };
            auto z_term = [&](int zz) {
                const float pz = s.bmin[2] + ((float)(z0 + zz) + 0.5f) * s.vs[2] + bp.sample.z;
                return m8 * pz;
            };
            // ParallelBeamsBackProjection.cl:26-52 for one sample
            auto update = [&](float xt, float zt, float value) {
                float u = xt + zt;
                u = u + m12 * 1.0f;
                float bu;
                const float tx = cl_fract(u - 0.5f, &bu);
                const int i = padded_index((int)bu, nu);
                const float L0 = lerp_pair(l0[i], tx), L1 = lerp_pair(l1[i], tx);
                const float L = L0 + (L1 - L0) * ty;
                const float R0 = lerp_pair(r0[i], tx), R1 = lerp_pair(r1[i], tx);
                const float R = R0 + (R1 - R0) * ty;
                const float wgt = chord / L;
                const float contrib = R * lam * wgt;
                const float sum = 0.0f + contrib;             // sum over the one sample
                value = (L > kEpsilon) ? value + sum : value; // += 1 * sum / 1 when used
                return S::round(value);
            };
            const int nrow = z1 - z0;
            if (kFusedThreads % gpr == 0) {
                // every thread keeps the same x group in every band row
                const int x0 = (tid % gpr) * G, dz = kFusedThreads / gpr;
                float xt[G];
#pragma unroll
                for (int i = 0; i < G; ++i) xt[i] = x_term(x0 + i);
#pragma unroll 1
                for (int zz = tid / gpr; zz < nrow; zz += dz) {
                    const float zt = z_term(zz);
                    float out[G];
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        out[i] = update(xt[i], zt, band[zz * nx + x0 + i]);
                        band[zz * nx + x0 + i] = out[i];
                    }
                    if (q == nbp - 1)
                        IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)x0, out);
                }
            } else {
#pragma unroll 1
                for (int gi = tid; gi < ng; gi += kFusedThreads) {
                    const int zz = gi / gpr, x0 = (gi - zz * gpr) * G;
                    const float zt = z_term(zz);
                    float out[G];
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        out[i] = update(x_term(x0 + i), zt, band[zz * nx + x0 + i]);
                        band[zz * nx + x0 + i] = out[i];
                    }
                    if (q == nbp - 1)
                        IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)x0, out);
                }
            }
            if (nbp > 1) __syncthreads();
        }
        if (nbp == 1) __syncthreads();

        // 3. forward projections of the current block through this band
        if (nfp > 0) {
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int id = k * kFusedThreads + tid;
                if (id < nrays) plane_advance<kCap>(ray[k], band, nx, nz, z0, z1);
            }
        }
        __syncthreads();
    }

    // ---- residual (+ long-object compensation) per pixel ----
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int id = k * kFusedThreads + tid;
        if (id < nrays) {
            const int q = id / rays_per_proj;
            const int rem = id - q * rays_per_proj;
            const int w = row_lo + rem / nu;
            const int u = rem - (rem / nu) * nu;
            const FusedFp& f = pass->fp[q];
            float acc = 0.0f, accl = 0.0f;
            acc += ray[k].acc;
            accl += ray[k].accl;
            float p = 0.0f + acc / 1.0f;
            float l = 0.0f + accl / 1.0f;
            if (f.use_loc) {
                p = l > kEpsilon ? p * f.loc_c / l : 0.0f;
                l = f.loc_c;
            }
            const size_t i = (size_t)w * nu + u;
            f.residual[i] = f.measured[i] - p;
            f.length[i] = l;
        }
    }

    // ---- rows whose rays miss the volume: zero line integral and length ----
    for (int d = j; d < fv.ndead; d += fv.ny) {
        const int w = fv.dead_rows[d];
        for (int q = 0; q < nfp; ++q) {
            const FusedFp& f = pass->fp[q];
            for (int u = tid; u < nu; u += kFusedThreads) {
                float p = 0.0f + 0.0f / 1.0f;
                float l = 0.0f + 0.0f / 1.0f;
                if (f.use_loc) {
                    p = l > kEpsilon ? p * f.loc_c / l : 0.0f;
                    l = f.loc_c;
                }
                const size_t i = (size_t)w * nu + u;
                f.residual[i] = f.measured[i] - p;
                f.length[i] = l;
            }
        }
    }
}

template <typename S, bool kVec, bool kCap>
void launch_fused_cap(int rpt, const FusedVolume& fv, const FusedPass* pass, size_t lds, hipStream_t stream)
{
    const dim3 grid(fv.ny), block(kFusedThreads);
    switch (rpt) {
        case 1: fused_plane_kernel<S, kVec, kCap, 1><<<grid, block, lds, stream>>>(fv, pass); break;
        case 2: fused_plane_kernel<S, kVec, kCap, 2><<<grid, block, lds, stream>>>(fv, pass); break;
        case 4: fused_plane_kernel<S, kVec, kCap, 4><<<grid, block, lds, stream>>>(fv, pass); break;
        default: fused_plane_kernel<S, kVec, kCap, 8><<<grid, block, lds, stream>>>(fv, pass); break;
    }
}

template <typename S>
void launch_fused_typed(int rpt, const FusedVolume& fv, const FusedPass* pass, size_t lds, hipStream_t stream)
{
    const bool vec = fv.nx % 4 == 0;
    const bool cap = fv.nx + fv.nz > 2049;
    if (vec && !cap) launch_fused_cap<S, true, false>(rpt, fv, pass, lds, stream);
    else if (vec) launch_fused_cap<S, true, true>(rpt, fv, pass, lds, stream);
    else if (!cap) launch_fused_cap<S, false, false>(rpt, fv, pass, lds, stream);
    else launch_fused_cap<S, false, true>(rpt, fv, pass, lds, stream);
}

}  // namespace

void launch_fused(etn_voxel_type vt, int rpt, const FusedVolume& fv, const FusedPass* pass, size_t lds,
                  hipStream_t stream)
{
    if (vt == ETN_VOXEL_FLOAT32) launch_fused_typed<F32>(rpt, fv, pass, lds, stream);
    else if (vt == ETN_VOXEL_FLOAT16) launch_fused_typed<F16>(rpt, fv, pass, lds, stream);
    else launch_fused_typed<BF16>(rpt, fv, pass, lds, stream);
    hip_check(hipGetLastError(), "fused plane kernel launch");
}

}  // namespace dev

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

class FusedPlaneEngineImpl final : public FusedPlaneEngine {
public:
    FusedPlaneEngineImpl(BlockIterativeReconstructionOperator& op, const etn_reconstruction_desc& d,
                         std::vector<int> row_lo, std::vector<int> row_hi, std::vector<int> dead_rows,
                         int rows_per_plane, hipStream_t stream);
    void enqueueIteration(hipStream_t stream) override;
    void check(hipStream_t stream) override;
    double timeKernels(hipStream_t stream, uint32_t* launches) override;

private:
    void build_passes();
    BlockIterativeReconstructionOperator& op_;
    etn_voxel_type vt_;
    int rows_per_plane_;
    int nu_, nv_;
    size_t pixels_;
    float lambda_;
    bool use_loc_;
    int max_fp_;
    dev::FusedVolume fv_;
    size_t lds_bytes_;
    DeviceBuffer rows_, diag_, images_, passes_dev_;
    std::vector<dev::FusedPass> passes_;
    std::vector<int> rpt_;
};

int select_rpt(int rays)
{
    int rpt = 1;
    while (rpt * dev::kFusedThreads < rays && rpt < dev::kMaxRaysPerThread) rpt *= 2;
    return rpt;
}

}  // namespace

std::unique_ptr<FusedPlaneEngine> FusedPlaneEngine::create(BlockIterativeReconstructionOperator& op,
                                                           const etn_reconstruction_desc& d, hipStream_t stream)
{
    const VolumeProperties& vp = op.getReconstructedVolume()->properties();
    if (vp.subvolume_count() != 1) return nullptr;
    if ((d.fp_samples ? d.fp_samples : 1) != 1 || (d.bp_samples ? d.bp_samples : 1) != 1) return nullptr;
    if (d.mask) return nullptr;
    const int nx = (int)vp.nx(), ny = (int)vp.ny(), nz = (int)vp.nz();
    const int nu = (int)d.base_geometry.resolution[0], nv = (int)d.base_geometry.resolution[1];
    if (8 * (nu + 2 * dev::kRowPad) + nx + 1 > dev::kLdsFloats) return nullptr;   // a z row and the detector rows fit in LDS

    // every projection: rays in y planes, constant detector row per plane
    std::vector<int> plane_of_row;
    for (uint32_t b = 0; b < op.GetBlockCount(); ++b)
        for (uint32_t p = 0; p < op.getBlockSize(b); ++p) {
            const ParallelScannerGeometry& g = op.geometry(op.getPhysicalProjectionIndex(b, p));
            const Vec3f r = g.ray_direction();
            if (r.y != 0.0f || g.horizontal_pitch().y != 0.0f) return nullptr;
            // nonzero x/z components must have finite reciprocals (the plane
            // DDA relies on every res value staying at or below FLT_MAX)
            for (float c : {r.x, r.z})
                if (c != 0.0f && std::fabs(c) < 1e-30f) return nullptr;
            const Matrix4x4 m = g.projection_matrix();
            const float* mv = m.data();
            if (mv[1] != 0.0f || mv[9] != 0.0f || mv[3] != 0.0f || mv[7] != 0.0f || mv[11] != 0.0f || mv[15] != 1.0f)
                return nullptr;
            // plane of each detector row, with the FP's own arithmetic (samples = 1);
            // -1 for rows whose sources lie outside the box in y by a margin that
            // makes every ray's y slab interval empty (t_out <= -2^-20 * FLT_MAX)
            std::vector<int> map(nv);
            const SubVolume& sub = vp.subvolume(0);
            const float ivs = 1.0f / vp.voxel_size().y;
            const float margin = 1.0f / 1048576.0f;
            for (int w = 0; w < nv; ++w) {
                const float fx = (float)0 + 0.5f + 1.0f * 0.0f;
                const float fy = (float)w + 0.5f + 1.0f * 0.0f;
                const float ru = g.horizontal_pitch().y * fx, rv = g.vertical_pitch().y * fy;
                float y = g.source_base().y + ru + rv;
                if (sub.bmin.y - y >= margin || y - sub.bmax.y >= margin) {
                    map[w] = -1;
                    continue;
                }
                if (y < sub.bmin.y || y > sub.bmax.y) return nullptr;   // too close to call: generic kernels
                y = std::fmin(std::fmax(y, sub.bmin.y), sub.bmax.y);
                int jj = (int)((y - sub.bmin.y) * ivs);
                jj = std::max(jj, 0);
                jj = std::min(jj, ny - 1);
                map[w] = jj;
            }
            if (plane_of_row.empty()) plane_of_row = map;
            else if (plane_of_row != map) return nullptr;
        }
    std::vector<int> lo(ny, 0), hi(ny, 0), dead;
    std::vector<int> seen(ny, 0);
    int prev = -1;
    for (int w = 0; w < nv; ++w) {
        const int jj = plane_of_row[w];
        if (jj < 0) {
            dead.push_back(w);
            continue;
        }
        if (jj != prev) {
            if (seen[jj]) return nullptr;        // rows of a plane must be contiguous
            seen[jj] = 1;
            lo[jj] = w;
            prev = jj;
        }
        hi[jj] = w + 1;
    }
    int rows_per_plane = 0;
    for (int jj = 0; jj < ny; ++jj) rows_per_plane = std::max(rows_per_plane, hi[jj] - lo[jj]);
    if (rows_per_plane * nu > dev::kFusedThreads * dev::kMaxRaysPerThread) return nullptr;
    (void)nz;
    return std::make_unique<FusedPlaneEngineImpl>(op, d, lo, hi, dead, rows_per_plane, stream);
}

FusedPlaneEngineImpl::FusedPlaneEngineImpl(BlockIterativeReconstructionOperator& op, const etn_reconstruction_desc& d,
                                           std::vector<int> row_lo, std::vector<int> row_hi,
                                           std::vector<int> dead_rows, int rows_per_plane, hipStream_t stream)
    : op_(op), vt_(d.voxel_type), rows_per_plane_(rows_per_plane), nu_((int)d.base_geometry.resolution[0]),
      nv_((int)d.base_geometry.resolution[1]), lambda_(d.lambda), use_loc_(d.use_long_object_compensation != 0)
{
    pixels_ = (size_t)nu_ * nv_;
    GpuVolume* vol = op.getReconstructedVolume();
    const VolumeProperties& vp = vol->properties();
    const int ny = (int)vp.ny();
    std::vector<int> table(row_lo);
    table.insert(table.end(), row_hi.begin(), row_hi.end());
    table.insert(table.end(), dead_rows.begin(), dead_rows.end());
    rows_.allocate(sizeof(int) * table.size());
    ETN_HIP(hipMemcpy(rows_.get(), table.data(), sizeof(int) * table.size(), hipMemcpyHostToDevice));
    diag_.allocate(sizeof(int));
    ETN_HIP(hipMemsetAsync(diag_.get(), 0, sizeof(int), stream));

    fv_.data = vol->data();
    fv_.slab = make_slab_params(vp, 0);
    fv_.nx = (int)vp.nx();
    fv_.ny = ny;
    fv_.nz = (int)vp.nz();
    fv_.nu = nu_;
    fv_.nv = nv_;
    fv_.row_lo = rows_.as<int>();
    fv_.row_hi = rows_.as<int>() + ny;
    fv_.dead_rows = rows_.as<int>() + 2 * ny;
    fv_.ndead = (int)dead_rows.size();
    fv_.diag = diag_.as<int>();
    // 64 KiB of LDS per workgroup: 4 padded detector rows, the rest a band of z rows
    const int row_floats = 8 * (nu_ + 2 * dev::kRowPad);   // 4 rows of (value, difference)
    fv_.band = std::max(1, std::min(fv_.nz, (dev::kLdsFloats - row_floats - 1) / fv_.nx));
    lds_bytes_ = sizeof(float) * ((((size_t)fv_.band * fv_.nx + 1) & ~(size_t)1) + (size_t)row_floats);
    max_fp_ = std::max(1, std::min(dev::kMaxFusedFp,
                                   dev::kFusedThreads * dev::kMaxRaysPerThread / std::max(1, rows_per_plane_ * nu_)));
    build_passes();
}

void FusedPlaneEngineImpl::build_passes()
{
    passes_.clear();
    rpt_.clear();
    const uint32_t nb = op_.GetBlockCount();
    uint32_t bmax = 0;
    for (uint32_t b = 0; b < nb; ++b) bmax = std::max(bmax, op_.getBlockSize(b));
    // residual + length images per block slot, ping-pong between consecutive blocks
    images_.allocate(sizeof(float) * pixels_ * 2 * 2 * bmax);
    auto image = [&](uint32_t parity, uint32_t slot, int which) {
        return images_.as<float>() + pixels_ * ((((size_t)parity * bmax + slot) * 2) + which);
    };
    const VolumeProperties& vp = op_.getReconstructedVolume()->properties();

    std::vector<dev::FusedBp> pending;   // back projections of the previous block
    auto flush_bp_only = [&](size_t keep) {
        // emit BP-only passes until at most `keep` back projections are pending
        while (pending.size() > keep) {
            dev::FusedPass p{};
            p.zdir = 1;
            const size_t n = std::min(pending.size() - keep, (size_t)dev::kMaxFusedBp);
            for (size_t i = 0; i < n; ++i) p.bp[p.nbp++] = pending[i];
            pending.erase(pending.begin(), pending.begin() + n);
            passes_.push_back(p);
        }
    };
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t bs = op_.getBlockSize(b);
        const uint32_t parity = b & 1u;
        // long-object compensation uses the block's last geometry (compensateLongObject)
        const ParallelScannerGeometry& last = op_.geometry(op_.getPhysicalProjectionIndex(b, bs - 1));
        const float loc_c = use_loc_ ? last.traversal_length_through(vp.bbox_min(), vp.bbox_max()) : 0.0f;
        std::vector<dev::FusedFp> neg, pos;
        for (uint32_t p = 0; p < bs; ++p) {
            const uint32_t phys = op_.getPhysicalProjectionIndex(b, p);
            const ParallelScannerGeometry& g = op_.geometry(phys);
            dev::FusedFp f{};
            g.source_base().store(f.src);
            g.horizontal_pitch().store(f.h);
            g.vertical_pitch().store(f.v);
            g.ray_direction().store(f.ray);
            f.measured = op_.realProjection(phys);
            f.residual = image(parity, p, 0);
            f.length = image(parity, p, 1);
            f.loc_c = loc_c;
            f.use_loc = use_loc_ ? 1 : 0;
            (f.ray[2] > 0.0f ? pos : neg).push_back(f);
        }
        // FP groups: rays moving towards -z (or staying in z) and towards +z
        std::vector<std::pair<int, std::vector<dev::FusedFp>>> groups;
        for (auto* grp : {&neg, &pos}) {
            const int zdir = grp == &pos ? 1 : -1;
            for (size_t i = 0; i < grp->size(); i += (size_t)max_fp_)
                groups.push_back({zdir, std::vector<dev::FusedFp>(grp->begin() + i,
                                                                  grp->begin() + std::min(grp->size(), i + (size_t)max_fp_))});
        }
        flush_bp_only(dev::kMaxFusedBp);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            dev::FusedPass p{};
            p.zdir = groups[gi].first;
            for (const auto& f : groups[gi].second) p.fp[p.nfp++] = f;
            if (gi == 0)
                for (const auto& e : pending) p.bp[p.nbp++] = e;
            passes_.push_back(p);
        }
        pending.clear();
        for (uint32_t p = 0; p < bs; ++p) {
            const ParallelScannerGeometry& g = op_.geometry(op_.getPhysicalProjectionIndex(b, p));
            dev::FusedBp e{};
            std::memcpy(e.m, g.projection_matrix().data(), sizeof(e.m));
            e.lambda = lambda_ / (float)bs;
            e.residual = image(parity, p, 0);
            e.length = image(parity, p, 1);
            const std::vector<VoxelSample> smp = bp_sample_pattern(g, vp.voxel_size(), 1);
            e.sample = make_float4(smp[0].dx, smp[0].dy, smp[0].dz, smp[0].length);
            pending.push_back(e);
        }
    }
    flush_bp_only(0);
    for (const auto& p : passes_) rpt_.push_back(select_rpt(p.nfp * rows_per_plane_ * nu_));
    passes_dev_.allocate(sizeof(dev::FusedPass) * passes_.size());
    ETN_HIP(hipMemcpy(passes_dev_.get(), passes_.data(), sizeof(dev::FusedPass) * passes_.size(),
                      hipMemcpyHostToDevice));
}

void FusedPlaneEngineImpl::enqueueIteration(hipStream_t stream)
{
    const dev::FusedPass* base = passes_dev_.as<dev::FusedPass>();
    for (size_t i = 0; i < passes_.size(); ++i) dev::launch_fused(vt_, rpt_[i], fv_, base + i, lds_bytes_, stream);
}

void FusedPlaneEngineImpl::check(hipStream_t stream)
{
    int bad = 0;
    ETN_HIP(hipMemcpyAsync(&bad, diag_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
    ETN_HIP(hipStreamSynchronize(stream));
    if (bad) throw Error(ETN_ERROR_INTERNAL, "fused plane kernel found rays outside their plane");
}

double FusedPlaneEngineImpl::timeKernels(hipStream_t stream, uint32_t* launches)
{
    // one eager iteration with an event pair around every launch; the average
    // over the passes that carry both a BP and an FP (the steady state)
    std::vector<hipEvent_t> ev(2 * passes_.size());
    for (auto& e : ev) ETN_HIP(hipEventCreate(&e));
    const dev::FusedPass* base = passes_dev_.as<dev::FusedPass>();
    for (size_t i = 0; i < passes_.size(); ++i) {
        ETN_HIP(hipEventRecord(ev[2 * i], stream));
        dev::launch_fused(vt_, rpt_[i], fv_, base + i, lds_bytes_, stream);
        ETN_HIP(hipEventRecord(ev[2 * i + 1], stream));
    }
    ETN_HIP(hipStreamSynchronize(stream));
    double total = 0.0;
    uint32_t n = 0;
    for (size_t i = 0; i < passes_.size(); ++i) {
        float ms = 0.0f;
        ETN_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        if (passes_[i].nfp > 0 && passes_[i].nbp > 0) {
            total += ms;
            ++n;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (launches) *launches = n;
    return n ? total / n : 0.0;
}

}  // namespace etn
