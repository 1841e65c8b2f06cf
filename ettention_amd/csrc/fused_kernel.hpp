// Fused plane kernels for single-axis tilt geometries (rays with no y
// component, detector rows of constant y).
//
// Every ray then stays in one volume plane y = j, and the back projection of
// a voxel in plane j reads detector rows that are constant across the plane.
// One workgroup owns a group of consecutive planes (their x rows are
// contiguous in HBM for every z).  It streams the planes through LDS in bands
// of z rows, in the direction the rays travel in z, and for each band
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
//
// A launch serves many independent reconstructions at once: its workgroups are
// (member, plane group) items, and each member contributes its current pass.
#pragma once

#include <cstdint>
#include <type_traits>

#include "device_math.hpp"
#include "fused.hpp"

namespace etn {
namespace dev {

constexpr int kMaxFusedFp = 16;      // forward projections per pass
constexpr int kMaxFusedBp = 8;       // back projections per pass
constexpr int kMaxRaysPerThread = 8;
constexpr int kMaxPlanes = 8;        // planes per workgroup
constexpr int kMaxFusedFpPerAxis = 4;   // forward sub-rays per pixel axis (oversampled kernels)
constexpr int kMaxFusedBpSamples = 16;  // back projection samples per voxel (oversampled kernels)
constexpr int kRowPad = 2;           // zero taps on each side of a staged detector row

// Workgroup shapes.  512 threads run two workgroups per CU; 256 threads with
// 40 KiB run four, whose independent barrier phases overlap better when a pass
// carries many projections, at half the band.
// kBand: band voxels; the vector path prefetches kBand / (4 kThreads) groups
// of 4 per thread in registers: 6 for f32 (24 VGPRs); 16-bit storage holds a
// group in 2 VGPRs, so its 512-thread shape takes 8 groups (taller bands,
// fewer barriers per sweep) in an 80 KiB array (gfx950 gives a workgroup up to
// 160 KiB; two of them fill a CU).  With a prior-knowledge mask the band holds
// kBand / 2 voxels and the voxels' mask weights follow at LDS offset kBand / 2
// (floats).  kBytes: bytes of a stored voxel.
template <int kThreads, int kBytes>
struct FusedShape;
template <>
struct FusedShape<512, 4> {
    static constexpr int kLds = 16384;
    static constexpr int kBand = 12288;
};
template <>
struct FusedShape<512, 2> {
    static constexpr int kLds = 20480;
    static constexpr int kBand = 16384;
};
template <int kBytes>
struct FusedShape<256, kBytes> {
    static constexpr int kLds = 10240;
    static constexpr int kBand = 6144;
};
// The register back projection with every record set of a pass resident in
// LDS (512 threads only): kRes == 2 takes a 160 KiB array, so one workgroup
// per CU (the 512-thread kernels with 4 or 8 rays per thread hold only one
// anyway, by their registers), with the band of the 512-thread shape;
// kRes == 4 takes an 80 KiB array with bands of at most 4096 voxels (the host
// passes a band height that fits) and 128 VGPRs, so that two workgroups share
// a CU.  kRes == 3 keeps the back projection on the band in LDS but takes the
// 160 KiB array too, for launches whose registers hold one 512-thread
// workgroup per CU anyway (4 or 8 rays per thread): wide volumes, whose
// detector records fill most of the 64 KiB shape, then get taller bands (the
// host passes the height).
constexpr int kResLds = 40960;
constexpr int kRes2Lds = 20480;
constexpr int kRes2Band = 4096;
template <int kThreads, int kBytes, int kRes>
struct FusedShapeOf {
    using type = FusedShape<kThreads, kBytes>;
};
template <int kBytes>
struct FusedShapeOf<512, kBytes, 2> {
    struct type {
        static constexpr int kLds = kResLds;
        static constexpr int kBand = FusedShape<512, kBytes>::kBand;
    };
};
template <int kBytes>
struct FusedShapeOf<512, kBytes, 3> {
    struct type {
        static constexpr int kLds = kResLds;
        static constexpr int kBand = FusedShape<512, kBytes>::kBand;
    };
};
template <int kBytes>
struct FusedShapeOf<512, kBytes, 4> {
    struct type {
        static constexpr int kLds = kRes2Lds;
        static constexpr int kBand = kRes2Band;
    };
};
constexpr int kMaxThreads = 512;
__host__ __device__ inline int shape_lds_floats(int threads, int vbytes)
{
    return threads == 256 ? FusedShape<256, 4>::kLds : vbytes == 2 ? FusedShape<512, 2>::kLds : FusedShape<512, 4>::kLds;
}
__host__ __device__ inline int shape_band_floats(int threads, bool mask, int vbytes)
{
    const int b = threads == 256 ? FusedShape<256, 4>::kBand
                                 : vbytes == 2 ? FusedShape<512, 2>::kBand : FusedShape<512, 4>::kBand;
    return mask ? b / 2 : b;
}

struct FusedFp {
    float src[3], h[3], v[3], ray[3];
    const void* measured;        // nv x nu, in measured_type (f32, or f16/bf16 stored projections)
    int measured_type;
    float* residual;             // nv x nu, written: measured - line integral (raw: the line integral)
    float* length;               // nv x nu, written
    float loc_c;
    int use_loc;
    int raw;                     // write the virtual projection itself (residual norm passes)
};

struct FusedBp {
    float m[16];
    float lambda;                // lambda / block size
    const float* residual;       // nv x nu
    const float* length;         // nv x nu
    float4 sample;               // single BP sample: offset xyz, in-voxel chord w
    // oversampled kernels: the per_dir x per_dir sample pattern, y-major (the
    // per_dir samples of one pattern row share a y offset)
    int nsamples, per_dir;
    float4 samples[kMaxFusedBpSamples];
    uint32_t sr_salt;            // stochastic-rounding kernels: per-projection salt of the random bits
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
    int planes;                  // consecutive y planes per workgroup (divides ny)
    int band;                    // z rows per LDS band
    int rows_per_plane;          // detector rows whose rays lie in a plane, maximum over planes
    int row_len;                 // staged columns per detector row, padding included
    int pad_l;                   // zero columns staged left of detector column 0
    const int* row_lo;           // per plane: first detector row whose rays lie in it
    const int* row_hi;           // per plane: one past the last such row
    const int* dead_rows;        // detector rows whose rays all miss the volume in y
    int ndead;
    const uint8_t* mask;         // prior-knowledge mask (complete volume, 255 = weight 1) or null: weight 1
    int fp_n;                    // oversampled kernels: forward sub-rays per pixel axis
    int fp_shift;                // log2 of the ray slots per pixel (a power of two >= fp_n: slot si traces sub-ray (si, 0))
    int rowsets;                 // oversampled kernels: staged row pairs per plane (BP pattern rows)
    int* diag;                   // counts live rays found outside their plane (must stay 0)
};

struct FusedWork {
    int member;                  // index into the launch's volume and pass tables
    int plane;                   // first plane of the workgroup's group
};

// LDS floats a workgroup of `threads` uses: the band (with a mask: the band,
// at offset shape_band_floats(threads, true, vbytes) the band's mask weights), then
// (16-byte aligned) the staged detector records of each plane and row set
// (row_floats per row set: 8 floats per padded column with exact arithmetic,
// 4 with fast; one row set per BP pattern row) and a (m4 py, ty) pair per
// plane and row set.
__host__ __device__ inline int fused_lds_floats(int planes, int band, int nx, int row_floats, bool mask, int threads,
                                                int rowsets, int vbytes)
{
    return (mask ? shape_band_floats(threads, true, vbytes) : 0) + ((band * planes * nx + 3) & ~3) +
           planes * rowsets * (row_floats + 2);
}

struct FusedLaunch {
    int threads;                 // workgroup shape: 512 or 256
    int rpt;
    int groups;
    const FusedVolume* vols;
    const FusedWork* work;
    const FusedPass* const* passes;
    size_t lds;
    bool sr;                     // stochastic rounding to 16-bit storage (fast kernels, no mask or oversampling)
    // back projections applied in registers as the band is loaded (fast
    // kernels, nx a multiple of 4, no oversampling, stochastic rounding or
    // step limit), the record sets of all of a pass's back projections staged
    // once per launch: 0 off, 1 in the shape's LDS array, 2 in the 160 KiB one,
    // 4 in the 80 KiB one with band_rows z rows per band; 3: off, with the
    // band in the 160 KiB array and band_rows z rows per band
    int res;
    int band_rows;               // > 0: z rows per band instead of each member's own
    // non-null: nothing is launched; the number of workgroups of the kernel
    // this launch selects that the device holds at once is stored here
    int* resident;
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
    int off;                     // LDS offset of the ray's plane inside a band row
    // without the step cap the band loop tracks the ray by its LDS byte offset
    // in the current band and by the x steps left before it leaves the volume
    int li, cx, dxb, dzb;
};

// Advances a ray while it is inside the volume and inside band [zlo, zhi).  A
// ray leaving the volume is finished; one leaving the band resumes in the next
// band with unchanged state.  kCap keeps the reference's 2048-step limit; it is
// dropped when no ray of the launch can visit more than 2040 voxels (the host
// bounds that by the rays' directions and nx + nz - 1, since every step moves
// vx or vz monotonically by one).
// kFast accumulates dt * value with a fused multiply-add (one rounding instead
// of two, as an OpenCL compiler may contract it).  kMask weights the ray
// length by the voxel's mask weight, stored kMaskOff floats after the voxel.
template <bool kCap, bool kFast, bool kMask, int kMaskOff>
__device__ __forceinline__ void plane_advance(PlaneRay& r, const float* band, int width, int nx, int nz, int zlo,
                                              int zhi)
{
    auto length_step = [](float accl, float dt, float mv) {
        if (!kMask) return accl + dt;                   // dt * 1.0f
        return kFast ? fmaf(dt, mv, accl) : accl + dt * mv;
    };
    if (!kCap) {
        // r.li is the byte offset of the ray's voxel in this band's LDS rows:
        // the voxel lies in the band exactly when 0 <= li < band bytes (its x
        // index is in range while x steps are left, so the row part stays
        // inside the ray's plane).  r.cx counts the x steps left inside the
        // volume.  The float operations are the loop below's: the step not
        // taken loses dt, |rx - rz| being exactly max - min.
        const unsigned lim = (unsigned)((zhi - zlo) * width) * 4u;
        int li = r.li, cx = r.cx;
        float rx = r.rx, rz = r.rz, acc = r.acc, accl = r.accl;
        const char* bytes = reinterpret_cast<const char*>(band);
        while (cx != 0 && (unsigned)li < lim) {
            const bool ix = rx <= rz;
            const float dt = ix ? rx : rz;
            const float value = *reinterpret_cast<const float*>(bytes + li);
            if (kFast) acc = fmaf(dt, value, acc);
            else acc += dt * value;
            accl = length_step(accl, dt, kMask ? *reinterpret_cast<const float*>(bytes + li + 4 * kMaskOff) : 1.0f);
            const float rest = fabsf(rx - rz);
            rx = ix ? r.stepx : rest;
            rz = ix ? rest : r.stepz;
            li += ix ? r.dxb : r.dzb;
            cx -= ix ? 1 : 0;
        }
        r.li = li;
        r.cx = cx;
        r.rx = rx;
        r.rz = rz;
        r.acc = acc;
        r.accl = accl;
        return;
    }
    if (!r.live) return;
    const unsigned h = (unsigned)(zhi - zlo);
    const int lz = r.sz * width;
    int li = (r.vz - zlo) * width + r.off + r.vx;
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
        if (kFast) acc = fmaf(dt, band[li], acc);
        else acc += dt * band[li];
        accl = length_step(accl, dt, kMask ? band[li + kMaskOff] : 1.0f);
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
// access granularity of the band transfers.  The band is read once and
// written once per sweep and not touched again within it, so the 4-voxel
// groups move with non-temporal hints: f32 loads and stores (512^3 SART:
// 621 -> 641 GUPS; either alone gained nothing), 16-bit loads only (724 ->
// 734 GUPS; non-temporal 16-bit stores lost 2%).
template <typename S, bool kVec>
struct GroupIo;
template <>
struct GroupIo<F32, true> {
    static constexpr int G = 4;
    using V = float4;
    typedef float f4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V load(const ETN_GLOBAL float* p, size_t i)
    {
        const f4 v = __builtin_nontemporal_load(reinterpret_cast<const ETN_GLOBAL f4*>(p + i));
        return make_float4(v.x, v.y, v.z, v.w);
    }
    static __device__ __forceinline__ void unpack(const V& v, float* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
    static __device__ __forceinline__ void store(ETN_GLOBAL float* p, size_t i, const float* f)
    {
        const f4 v = {f[0], f[1], f[2], f[3]};
        __builtin_nontemporal_store(v, reinterpret_cast<ETN_GLOBAL f4*>(p + i));
    }
};
template <typename S16>
struct GroupIo16 {   // 4 x 16-bit voxels as one 8-byte access
    static constexpr int G = 4;
    using V = uint2;
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ V load(const ETN_GLOBAL typename S16::T* p, size_t i)
    {
        const u2 v = __builtin_nontemporal_load(reinterpret_cast<const ETN_GLOBAL u2*>(p + i));
        return make_uint2(v.x, v.y);
    }
    static __device__ __forceinline__ void unpack(const V& v, float* o)
    {
        S16::unpack2(v.x, o[0], o[1]);
        S16::unpack2(v.y, o[2], o[3]);
    }
    static __device__ __forceinline__ void store(ETN_GLOBAL typename S16::T* p, size_t i, const float* f)
    {
        V v;
        v.x = S16::pack2(f[0], f[1]);
        v.y = S16::pack2(f[2], f[3]);
        *reinterpret_cast<ETN_GLOBAL uint2*>(p + i) = v;
    }
    // the group rounded to storage (round to nearest even, as S16::round does
    // value by value) in two packed conversions: the rounded values back in
    // f, the packed group returned for store_packed
    static __device__ __forceinline__ V round4(float* f)
    {
        V v;
        v.x = S16::pack2(f[0], f[1]);
        v.y = S16::pack2(f[2], f[3]);
        S16::unpack2(v.x, f[0], f[1]);
        S16::unpack2(v.y, f[2], f[3]);
        return v;
    }
    static __device__ __forceinline__ void store_packed(ETN_GLOBAL typename S16::T* p, size_t i, const V& v)
    {
        *reinterpret_cast<ETN_GLOBAL uint2*>(p + i) = v;
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
    static __device__ __forceinline__ V load(const ETN_GLOBAL typename S::T* p, size_t i) { return p[i]; }
    static __device__ __forceinline__ void unpack(const V& v, float* o) { o[0] = S::load(&v, 0); }
    static __device__ __forceinline__ void store(ETN_GLOBAL typename S::T* p, size_t i, const float* f)
    {
        S::store(p, i, f[0]);
    }
};

// LinearInterpolation (ImageHandler.hl:20-29) on staged rows with kRowPad zeros on
// each side, stored as (v[i], v[i+1] - v[i]) pairs: clamping the base index to
// [-2, n] reads only zeros wherever the reference's range checks return 0, and
// the stored difference is the subtraction mix() performs, so the value is the same.
// Column i of the two residual rows and the two length rows forms one 32-byte
// record {r0, dr0, r1, dr1}, {l0, dl0, l1, dl1}: one address, two 16-byte reads.
__device__ __forceinline__ int padded_index(int i0, int n) { return min(max(i0, -kRowPad), n) + kRowPad; }
__device__ __forceinline__ float lerp_pair(float a, float d, float t) { return a + d * t; }

// One single-sample back projection update of a voxel with fast arithmetic
// (ParallelBeamsBackProjection.cl:26-52): xt + zt = m0 px + m4 py + m8 pz, u
// in the reference's order (u is a few hundred pixels, so a reordered sum
// would move the weights by ~1e-4).  The bilinear base column and weight come
// from one floor-convert and one fract (fract clamps below 1: a tx that
// u - floor(u) would round up to 1 stays 1 - 2^-24, the limit the
// interpolation approaches either way).  pr: the plane's staged records from
// column -pad_l on, padded (on the host) beyond every column the member's
// projections reach, so the index needs no clamp, with the residual pair
// scaled by lambda times the in-voxel chord at staging.  mv: the mask weight;
// vid: the voxel's index in the volume (stochastic rounding's random bits).
__device__ __forceinline__ int floor_to_int(float u)
{
    int i;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(u));
    return i;
}
template <typename S, bool kMask, bool kSR, bool kRound = true>
__device__ __forceinline__ float bp_update_fast(const float4* pr, float xt, float zt, float m12, float value, float mv,
                                                uint32_t vid, uint32_t salt)
{
    float u = xt + zt;
    u = u + m12 * 1.0f;
    u = u - 0.5f;
    const float tx = __builtin_amdgcn_fractf(u);
    const float4 c = pr[floor_to_int(u)];
    const float Rv = fmaf(c.y, tx, c.x), L = fmaf(c.w, tx, c.z);
    const float contrib = Rv * __builtin_amdgcn_rcpf(L);
    const float next = kMask ? fmaf(mv, contrib, value) : value + contrib;
    value = (L > kEpsilon) ? next : value;
    if constexpr (!kRound) return value;   // the caller rounds (round4)
    if constexpr (kSR) return S::round_stochastic(value, sr_hash(vid, salt, value));
    return S::round(value);
}

// One workgroup per (member, plane group) work item.  `passes[m]` is member m's
// pass of this launch; members without one get no work items.
//
// kFast keeps every operation of the reference's algorithm but not its
// rounding order: the bilinear lookup becomes A + B tx with A and B combined
// per detector column at staging time (exactly the same function of tx and ty,
// rounded differently), the weight chord / L uses the hardware reciprocal
// (1 ulp; OpenCL allows 2.5 ulp for a division), and the ray sums use fused
// multiply-adds.  Without kFast the kernel performs the generic kernels'
// float operations one for one.
// kMask applies the prior-knowledge mask (ImageHandler.hl:85-89): the ray
// length sums dt * weight, the update is weight * sum.  A member without a
// mask gets weight 1 (every update then equals the unmasked kernel's).
// kOver oversamples: fp_n x fp_n sub-rays per detector pixel
// (ForwardProjectionKernel.cpp:65-68), whose sums the pixel averages, and a
// per_dir x per_dir sample pattern per voxel (ParallelBeamsBackProjectionKernel
// .cpp:13-103), each pattern row with its own staged row pair.  Sub-rays (si,
// sj) of one si differ only in y (the detector's v axis is y, checked on the
// host) and lie in the same plane, so they take the same x-z steps through the
// same voxels and give bit-identical sums: the kernel traces sub-ray (si, 0)
// and adds its sums fp_n times, in the generic kernel's (si, sj) order.
// With at most two rays per thread the kernel fits 128 VGPRs: four waves per
// SIMD, two 512-thread workgroups per CU (as LDS also allows).
// kRes != 0 (fast arithmetic, nx a multiple of 4, no oversampling or
// stochastic rounding) applies the back projections to each group of 4 voxels
// in registers as the band arrives from HBM, all of a pass's back projections
// in block order, before the group goes to LDS: no LDS read and write of the
// band per back projection and no barrier between them.  The record sets of
// every back projection of the pass are staged once per launch (kRes == 2: in
// the 160 KiB array).  Every voxel sees the same chain of updates and
// roundings as on the LDS path, so the volumes are bit-identical.
template <typename S, bool kVec, bool kCap, bool kFast, bool kMask, bool kOver, int kThreads, int RPT, bool kSR,
          int kRes>
__global__ void __launch_bounds__(kThreads, kRes == 2 || kRes == 3 ? 2
                                            : (kRes == 4 || RPT <= 2 || (kOver && RPT == 4)) ? 4 : 1)
    fused_plane_kernel(const FusedVolume* __restrict__ vols, const FusedWork* __restrict__ work,
                       const FusedPass* const* __restrict__ passes, int band_rows)
{
    // kRes 1, 2, 4: the back projection in registers; 0 and 3: on the band in LDS
    constexpr bool kReg = kRes == 1 || kRes == 2 || kRes == 4;
    static_assert(!kReg || (kVec && kFast && !kOver && !kSR), "register back projection: fast vector kernels");
    static_assert((kRes != 2 && kRes != 3 && kRes != 4) || kThreads == 512, "160 and 80 KiB arrays: 512 threads");
    using IO = GroupIo<S, kVec>;
    using Shape = typename FusedShapeOf<kThreads, (int)sizeof(typename S::T), kRes>::type;
    constexpr int G = IO::G;
    constexpr int kMaskOff = Shape::kBand / 2;          // LDS floats from a voxel to its mask weight
    constexpr int kBandCap = kMask ? kMaskOff : Shape::kBand;
    constexpr int kGroups = kBandCap / (G * kThreads);  // band groups per thread (upper bound)
    // a fixed-size array puts the band at LDS address 0 (the host sizes every
    // tile to fit it; the VGPR budget limits the workgroups per CU as much)
    __shared__ __attribute__((aligned(16))) float lds[Shape::kLds];
    const FusedWork item = work[blockIdx.x];
    const ETN_GLOBAL FusedVolume& fv = global_ptr(vols)[item.member];
    const int fband = band_rows > 0 ? band_rows : fv.band;   // z rows per band
    const ETN_GLOBAL FusedPass* pass = global_ptr(global_ptr(passes)[item.member]);
    const int j = item.plane;                           // planes j .. j + np - 1
    const int tid = threadIdx.x;
    const int nx = fv.nx, nz = fv.nz, nu = fv.nu, np = fv.planes;
    const int width = np * nx;                          // one band row: the x rows of the np planes
    const int row_len = fv.row_len, pad_l = fv.pad_l;
    float* band = lds;                                  // band x width
    // per plane and row set row_len column records, 16-byte aligned after the
    // band, then per plane and row set (m4 py, ty) of the current back projection
    const int rowsets = kOver ? fv.rowsets : 1;
    float* mband = lds + kMaskOff;                      // kMask: the band's mask weights
    float4* recs = reinterpret_cast<float4*>(lds + (kMask ? kMaskOff : 0) + ((fband * width + 3) & ~3));
    constexpr int kRecStride = kFast ? 1 : 2;   // float4 per column record
    ETN_GLOBAL typename S::T* vol = global_ptr(static_cast<typename S::T*>(fv.data));
    const ETN_GLOBAL uint8_t* mask = kMask ? global_ptr(fv.mask) : nullptr;
    // MaskVolume weight of a byte, as the generic kernels compute it
    auto weight = [](uint32_t m) { return (float)m / 255.0f; };
    const size_t plane_stride = (size_t)nx * fv.ny;
    const size_t plane_base = (size_t)j * nx;
    const int gpr = width / G;                          // groups per band row

    const int nfp = pass->nfp, nbp = pass->nbp;
    // record sets staged at a time (set q: back projection q; the LDS path
    // restages set 0 per projection when a pass has several)
    const int nsets = kReg ? max(nbp, 1) : 1;
    const int set_len = np * rowsets * kRecStride * row_len;   // float4 per record set
    float2* pterm = reinterpret_cast<float2*>(recs + nsets * set_len);
    const int R = fv.rows_per_plane;
    const int fp_n = kOver ? fv.fp_n : 1;
    const int fp_shift = kOver ? fv.fp_shift : 0;
    const int rays_per_proj = (np * R * nu) << fp_shift;
    const int nrays = nfp * rays_per_proj;
    const SlabParams s = fv.slab;
    const ETN_GLOBAL int* row_lo = global_ptr(fv.row_lo);
    const ETN_GLOBAL int* row_hi = global_ptr(fv.row_hi);

    // ray slot id = (((q * np + p) * R + r) * nu + u) * 2^fp_shift + si:
    // projection q, plane j + p, its r-th detector row w, column u, sub-ray
    // column si; slots past a plane's rows or past fp_n stay idle
    auto decode = [&](int id, int& q, int& p, int& w, int& u, int& sub) {
        q = id / rays_per_proj;
        int rem = id - q * rays_per_proj;
        sub = rem & ((1 << fp_shift) - 1);
        rem >>= fp_shift;
        const int pr = rem / nu;
        u = rem - pr * nu;
        p = pr / R;
        w = row_lo[j + p] + (pr - p * R);
        return w < row_hi[j + p] && sub < fp_n;
    };

    const int nbands = (nz + fband - 1) / fband;

    // v coordinate of plane j + p under back projection bp for a sample with y
    // offset dy, constant over the plane (checked on the host): returns
    // detector row j0, sets the weight ty and the plane's voxel-centre y
    auto plane_v = [&](const ETN_GLOBAL FusedBp& bp, int p, float dy, float& py, float& ty) {
        py = s.bmin[1] + ((float)(j + p) + 0.5f) * s.vs[1];
        float vp = bp.m[1] * 0.0f + bp.m[5] * (py + dy);
        vp = vp + bp.m[9] * 0.0f;
        vp = vp + bp.m[13] * 1.0f;
        float base_row;
        ty = cl_fract(vp - 0.5f, &base_row);
        return (int)base_row;
    };
    auto stage_rows = [&](int q, int set) {
        const ETN_GLOBAL FusedBp& bp = pass->bp[q];
        float4* dst = recs + set * set_len;
        const ETN_GLOBAL float* res = global_ptr(bp.residual);
        const ETN_GLOBAL float* len = global_ptr(bp.length);
        // record e = (p * rowsets + ys) * row_len + k: plane p, pattern row ys, column k
        const int per_plane = rowsets * row_len;
        for (int e = tid; e < np * per_plane; e += kThreads) {
            const int p = e / per_plane, rem = e - p * per_plane, ys = rem / row_len;
            const int k = rem - ys * row_len, u = k - pad_l;
            const float dy = kOver ? bp.samples[ys * bp.per_dir].y : bp.sample.y;
            float py, ty;
            const int row = plane_v(bp, p, dy, py, ty);
            float v[8];   // (value, difference) of residual row, row + 1, length row, row + 1
            for (int rr = 0; rr < 2; ++rr) {
                const bool rok = row + rr >= 0 && row + rr < fv.nv;
                const bool ok0 = rok && u >= 0 && u < nu, ok1 = rok && u + 1 >= 0 && u + 1 < nu;
                const size_t at = (size_t)(row + rr) * nu + u;
                const float r0 = ok0 ? res[at] : 0.0f;
                const float r1 = ok1 ? res[at + 1] : 0.0f;
                const float l0 = ok0 ? len[at] : 0.0f;
                const float l1 = ok1 ? len[at + 1] : 0.0f;
                v[2 * rr] = r0;
                v[2 * rr + 1] = r1 - r0;
                v[4 + 2 * rr] = l0;
                v[4 + 2 * rr + 1] = l1 - l0;
            }
            if (kFast) {
                // R(tx) = A + B tx with A, B the row pair interpolated at ty; a
                // single sample's A and B carry the weight lambda * chord
                const float lc = kOver ? 1.0f : bp.lambda * bp.sample.w;
                dst[e] = make_float4(lc * fmaf(v[2] - v[0], ty, v[0]), lc * fmaf(v[3] - v[1], ty, v[1]),
                                     fmaf(v[6] - v[4], ty, v[4]), fmaf(v[7] - v[5], ty, v[5]));
            } else {
                dst[2 * e] = make_float4(v[0], v[1], v[2], v[3]);
                dst[2 * e + 1] = make_float4(v[4], v[5], v[6], v[7]);
            }
        }
        if (tid < np * rowsets) {
            const int p = tid / rowsets, ys = tid - p * rowsets;
            const float dy = kOver ? bp.samples[ys * bp.per_dir].y : bp.sample.y;
            float py, ty;
            plane_v(bp, p, dy, py, ty);
            pterm[set * np * rowsets + tid] = make_float2(bp.m[4] * (py + dy), ty);
        }
    };

    auto band_range = [&](int bi, int& z0, int& z1) {
        const int b = pass->zdir >= 0 ? bi : nbands - 1 - bi;
        z0 = b * fband;
        z1 = min(z0 + fband, nz);
    };

    // the vector path keeps the next band in registers while this band is
    // processed; the scalar path (nx not a multiple of 4) loads at commit time
    constexpr int kPre = kVec ? kGroups : 1;
    typename IO::V pre[kPre];
    uint32_t mpre[kMask ? kPre : 1];                    // the groups' 4 mask bytes
    // groups [k0, k1) of each thread's share of band [z0, z1)
    auto prefetch = [&](int z0, int z1, int k0, int k1) {
        if (!kVec) return;
        const int ng = (z1 - z0) * gpr;
#pragma unroll
        for (int k = 0; k < kPre; ++k) {
            if (k < k0 || k >= k1) continue;
            const int gi = k * kThreads + tid;
            if (gi < ng) {
                const int zz = gi / gpr, xg = gi - zz * gpr;
                const size_t at = (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xg * G;
                pre[k] = IO::load(vol, at);
                if (kMask) mpre[k] = mask ? *reinterpret_cast<const ETN_GLOBAL uint32_t*>(mask + at) : 0xffffffffu;
            }
        }
    };

    // with at most two rays per thread the first band's loads go out before
    // the rays are set up (more rays: their set-up would hold the loads'
    // registers too)
    int z0, z1;
    band_range(0, z0, z1);
    if (RPT <= 2) prefetch(z0, z1, 0, kPre);

    // the measured projection's value in its storage type (a uniform branch)
    auto measured_at = [](const ETN_GLOBAL FusedFp& f, size_t i) -> float {
        if (f.measured_type == ETN_VOXEL_FLOAT16) return F16::load(global_ptr(static_cast<const __half*>(f.measured)), i);
        if (f.measured_type == ETN_VOXEL_BFLOAT16)
            return BF16::load(global_ptr(static_cast<const uint16_t*>(f.measured)), i);
        return global_ptr(static_cast<const float*>(f.measured))[i];
    };

    // ---- ray set-up: dda_begin exactly as the generic FP kernel ----
    PlaneRay ray[RPT];
    // with at most two rays per thread each ray's measured pixel is loaded
    // here, so that the residual does not wait for it after the last band
    constexpr bool kEarly = RPT <= 2;
    [[maybe_unused]] float meas[kEarly ? RPT : 1];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int id = k * kThreads + tid;
        ray[k].live = 0;
        ray[k].acc = 0.0f;
        ray[k].accl = 0.0f;
        ray[k].off = 0;
        int q, p, w, u, sub;
        if (id < nrays && decode(id, q, p, w, u, sub)) {
            const ETN_GLOBAL FusedFp& f = pass->fp[q];
            if (kEarly && sub == 0 && !f.raw) meas[k] = measured_at(f, (size_t)w * nu + u);
            RayParams g;
            for (int a = 0; a < 3; ++a) {
                g.source[a] = f.src[a];
                g.h[a] = f.h[a];
                g.v[a] = f.v[a];
                g.ray[a] = f.ray[a];
            }
            const DdaConst c = dda_const(g.ray, s.vs);
            float src[3];
            if (kOver) {
                // the generic FP's sub-ray loop: sd = 1 / n, offset sd / 2; sub-ray (si, 0)
                const float sd = 1.0f / (float)fp_n;
                fp_ray_source(g, (uint32_t)u, (uint32_t)w, (uint32_t)sub, 0, sd, sd / 2.0f, src);
            } else {
                fp_ray_source(g, (uint32_t)u, (uint32_t)w, 0, 0, 1.0f, 0.5f, src);
            }
            DdaRay r;
            dda_begin(s, c, src, r);
            if (r.live && r.vox[1] != j + p) {
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
            ray[k].off = p * nx;
        }
    }

    if (!kCap) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            PlaneRay& r = ray[k];
            r.li = ((r.vz - z0) * width + r.off + r.vx) * 4;
            r.cx = !r.live ? 0 : r.sx > 0 ? nx - r.vx : r.sx < 0 ? r.vx + 1 : 0x7fffffff;
            r.dxb = 4 * r.sx;
            r.dzb = 4 * r.sz * width;
        }
    }
    if (RPT > 2) prefetch(z0, z1, 0, kPre);
    if constexpr (kReg) {
        // every back projection's records, before the first band needs them
        if (nbp > 0) {
            for (int q = 0; q < nbp; ++q) stage_rows(q, q);
            __syncthreads();
        }
    } else if (nbp == 1) {
        stage_rows(0, 0);            // made visible by the first band's barrier
    }
    // register back projection: the host launches it only where kThreads is a
    // multiple of the groups per band row, so that a thread's groups keep one
    // x in every band row; the voxel-centre x of the thread's 4 voxels, their
    // plane and first x; group k's band row rz0 + k rdz
    [[maybe_unused]] int rp = 0, rx0 = 0;
    [[maybe_unused]] const int rz0 = tid / gpr, rdz = kThreads / gpr;
    [[maybe_unused]] const float4* rrec = recs;   // the plane's records of set 0 from column -pad_l on
    [[maybe_unused]] float xc[kReg ? 4 : 1];
    if constexpr (kReg) {
        const int xw = (tid % gpr) * 4;
        if (np > 1) {
            rp = xw / nx;
            rx0 = xw - rp * nx;
        } else {
            rx0 = xw;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) xc[i] = s.bmin[0] + ((float)(rx0 + i) + 0.5f) * s.vs[0];
        rrec = recs + rp * row_len + pad_l;
    }

    for (int bi = 0; bi < nbands; ++bi) {
        band_range(bi, z0, z1);
        const int ng = (z1 - z0) * gpr;
        // the next band (loaded in two halves with f32 storage: see below)
        constexpr int kSplit = sizeof(typename S::T) == 4 ? (kPre + 1) / 2 : kPre;
        int nb0 = 0, nb1 = 0;
        if (bi + 1 < nbands) band_range(bi + 1, nb0, nb1);
        if constexpr (kReg) {
            // 1. the prefetched band with the previous block's back projections
            //    applied in registers, then to LDS and HBM
            float v[kPre][4];
            float mw[kMask ? kPre : 1][4];
#pragma unroll
            for (int k = 0; k < kPre; ++k) {
                const int gi = k * kThreads + tid;
                if (gi < ng) {
                    IO::unpack(pre[k], v[k]);
                    if (kMask) {
                        const uint32_t m = mpre[k];
                        mw[k][0] = weight(m & 255u);
                        mw[k][1] = weight((m >> 8) & 255u);
                        mw[k][2] = weight((m >> 16) & 255u);
                        mw[k][3] = weight(m >> 24);
                    }
                }
            }
            // voxel-centre z of group k's row (kept across the projections)
            float zc[kPre];
#pragma unroll
            for (int k = 0; k < kPre; ++k) zc[k] = s.bmin[2] + ((float)(z0 + rz0 + k * rdz) + 0.5f) * s.vs[2];
            if (bi + 1 < nbands) prefetch(nb0, nb1, 0, kSplit);
            // u = ((m0 px + m4 py) + m8 pz) + m12 w, w == 1 exactly (checked on the
            // host), px and pz the voxel centre plus the sample offset
#pragma unroll 1
            for (int q = 0; q < nbp; ++q) {
                const ETN_CONST FusedBp& bp = const_ptr(pass)->bp[q];
                const float m0 = bp.m[0], m8 = bp.m[8], m12 = bp.m[12];
                const float4 smp = bp.sample;
                const float2 pt = pterm[q * np + rp];
                // set q's records: their LDS byte offset as one VGPR, so that a
                // record address is a single shift-add of the column
                uint32_t prb = (uint32_t)(reinterpret_cast<const char*>(rrec + q * set_len) -
                                          reinterpret_cast<const char*>(lds));
                asm("" : "+v"(prb));
                const float4* pr = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(lds) + prb);
                float xt[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) xt[i] = m0 * (xc[i] + smp.x) + pt.x;
#pragma unroll
                for (int k = 0; k < kPre; ++k) {
                    if (k * kThreads + tid < ng) {
                        const float zt = m8 * (zc[k] + smp.z);
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            v[k][i] = bp_update_fast<S, kMask, false, false>(pr, xt[i], zt, m12, v[k][i],
                                                                             kMask ? mw[k][i] : 1.0f, 0u, 0u);
                        // rounded to storage after every projection
                        if constexpr (sizeof(typename S::T) == 2) IO::round4(v[k]);
                    }
                    // one group's record reads in flight at a time: hoisting all
                    // of them would take 4 VGPRs per voxel
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
#pragma unroll
            for (int k = 0; k < kPre; ++k) {
                if (k * kThreads + tid < ng) {
                    const int zz = rz0 + k * rdz, xw = (tid % gpr) * 4;
                    *reinterpret_cast<float4*>(band + zz * width + xw) = make_float4(v[k][0], v[k][1], v[k][2], v[k][3]);
                    if (kMask)
                        *reinterpret_cast<float4*>(mband + zz * width + xw) =
                            make_float4(mw[k][0], mw[k][1], mw[k][2], mw[k][3]);
                    if (nbp > 0) IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw, v[k]);
                }
            }
            __syncthreads();
        } else {
            // 1. prefetched band -> LDS
            if (kVec) {
#pragma unroll
                for (int k = 0; k < kPre; ++k) {
                    const int gi = k * kThreads + tid;
                    if (gi < ng) {
                        float v[G];
                        IO::unpack(pre[k], v);
                        const int zz = gi / gpr, xg = gi - zz * gpr;
                        // one 16-byte LDS store (width and xg * G are multiples of 4)
                        *reinterpret_cast<float4*>(band + zz * width + xg * G) = make_float4(v[0], v[1], v[2], v[3]);
                        if (kMask) {
                            const uint32_t m = mpre[k];
                            *reinterpret_cast<float4*>(mband + zz * width + xg * G) =
                                make_float4(weight(m & 255u), weight((m >> 8) & 255u), weight((m >> 16) & 255u),
                                            weight(m >> 24));
                        }
                    }
                }
            } else {
                for (int gi = tid; gi < ng; gi += kThreads) {
                    const int zz = gi / gpr, xg = gi - zz * gpr;
                    const size_t at = (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xg * G;
                    float v[G];
                    IO::unpack(IO::load(vol, at), v);
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        band[zz * width + xg * G + i] = v[i];
                        if (kMask) mband[zz * width + xg * G + i] = mask ? weight(mask[at + i]) : 1.0f;
                    }
                }
            }
            // the prefetch registers are free again: start fetching the next band
            // before waiting for the other waves (it only reads planes this
            // workgroup stores later).  With f32 storage half of the groups go
            // now and the other half before the forward projections, so that
            // memory requests stay in flight through both phases of the band (the
            // f32 launch is bound by how well memory overlaps the arithmetic:
            // 568 -> 612 GUPS at 512^3).  16-bit storage moves half the bytes and
            // is bound by VALU issue; it loads the whole band now (split, it lost 4%).
            if (bi + 1 < nbands) prefetch(nb0, nb1, 0, kSplit);
            // every thread back-projects exactly the groups it committed, so
            // with back projections the band needs no barrier here (the one
            // after them makes it visible to the forward projections); the
            // first band's makes the records staged before the loop visible
            if (bi == 0 || nbp == 0) __syncthreads();

            // 2. back projections of the previous block, in block order; each thread
            //    updates (and finally stores) the groups it staged
            for (int q = 0; q < nbp; ++q) {
                if (nbp > 1) {
                    stage_rows(q, 0);
                    __syncthreads();
                }
                // the pass lives in global memory: copy what the loops use into
                // registers (scalar loads: the next band's loads are in flight)
                const ETN_CONST FusedBp& bp = const_ptr(pass)->bp[q];
                const float m0 = bp.m[0], m8 = bp.m[8], m12 = bp.m[12], lam = bp.lambda;
                const float4 smp = bp.sample;
                const float chord = smp.w;
                const uint32_t salt = kSR ? bp.sr_salt : 0u;
                // u = ((m0 px + m4 py) + m8 pz) + m12 w, w == 1 exactly (checked on the host)
                auto x_term = [&](int x, float a4) {
                    const float px = s.bmin[0] + ((float)x + 0.5f) * s.vs[0] + smp.x;
                    return m0 * px + a4;
                };
                auto z_term = [&](int zz) {
                    const float pz = s.bmin[2] + ((float)(z0 + zz) + 0.5f) * s.vs[2] + smp.z;
                    return m8 * pz;
                };
                // ParallelBeamsBackProjection.cl:26-52 for one sample; pr: the plane's records
                // vid: the voxel's index in the volume (stochastic rounding's random bits)
                auto update = [&](const float4* pr, float ty, float xt, float zt, float value, float mv, uint32_t vid) {
                    if (kFast) return bp_update_fast<S, kMask, kSR>(pr + pad_l, xt, zt, m12, value, mv, vid, salt);
                    float u = xt + zt;
                    u = u + m12 * 1.0f;
                    float bu;
                    const float tx = cl_fract(u - 0.5f, &bu);
                    const int i = padded_index((int)bu, nu);
                    const float4 rr = pr[2 * i], ll = pr[2 * i + 1];
                    const float L0 = lerp_pair(ll.x, ll.y, tx), L1 = lerp_pair(ll.z, ll.w, tx);
                    const float L = L0 + (L1 - L0) * ty;
                    const float R0 = lerp_pair(rr.x, rr.y, tx), R1 = lerp_pair(rr.z, rr.w, tx);
                    const float Rv = R0 + (R1 - R0) * ty;
                    const float wgt = chord / L;
                    const float contrib = Rv * lam * wgt;
                    const float sum = 0.0f + contrib;             // sum over the one sample
                    const float add = kMask ? mv * sum : sum;     // weight * sum / 1
                    value = (L > kEpsilon) ? value + add : value; // used == 1 exactly when L > eps
                    return S::round(value);
                };
                const int nrow = z1 - z0;
                if constexpr (kOver) {
                    // the sample loop of ParallelBeamsBackProjection.cl:26-52 over the
                    // pattern; pattern row ys reads its own staged row pair
                    const int nsmp = bp.nsamples, pd = bp.per_dir;
#pragma unroll 1
                    for (int gi = tid; gi < ng; gi += kThreads) {
                        const int zz = gi / gpr, xw = (gi - zz * gpr) * G;
                        int p = 0, x0 = xw;
                        if (np > 1) {
                            p = xw / nx;
                            x0 = xw - p * nx;
                        }
                        const float cz = s.bmin[2] + ((float)(z0 + zz) + 0.5f) * s.vs[2];
                        // samples outer, the group's voxels inner: each sample's
                        // parameters and z term serve the G voxels, whose G
                        // independent chains overlap their record reads; every
                        // voxel still adds its samples in pattern order
                        float cx[G], sum[G];
                        int used[G];
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            cx[i] = s.bmin[0] + ((float)(x0 + i) + 0.5f) * s.vs[0];
                            sum[i] = 0.0f;
                            used[i] = 0;
                        }
#pragma unroll 1
                        for (int t = 0; t < nsmp; ++t) {
                            const float4 sm = bp.samples[t];
                            const int ys = t / pd;
                            const float2 pt = pterm[p * rowsets + ys];
                            const float zt = m8 * (cz + sm.z);
                            const float lw = lam * sm.w;
#pragma unroll
                            for (int i = 0; i < G; ++i) {
                                float u = m0 * (cx[i] + sm.x) + pt.x;
                                u = u + zt;
                                u = u + m12 * 1.0f;
                                float L, contrib;
                                if (kFast) {
                                    u = u - 0.5f;
                                    const float bu = floorf(u);
                                    const float tx = u - bu;
                                    const float4 c = recs[(p * rowsets + ys) * row_len + (int)bu + pad_l];
                                    const float Rv = fmaf(c.y, tx, c.x);
                                    L = fmaf(c.w, tx, c.z);
                                    contrib = Rv * lw * __builtin_amdgcn_rcpf(L);
                                } else {
                                    // bp_voxel_sum's bilinear lookups and weight, operation for operation
                                    float bu;
                                    const float tx = cl_fract(u - 0.5f, &bu);
                                    const int k = padded_index((int)bu, nu);
                                    const float4* pr = recs + (p * rowsets + ys) * 2 * row_len;
                                    const float4 rr = pr[2 * k], ll = pr[2 * k + 1];
                                    const float L0 = lerp_pair(ll.x, ll.y, tx), L1 = lerp_pair(ll.z, ll.w, tx);
                                    L = L0 + (L1 - L0) * pt.y;
                                    const float R0 = lerp_pair(rr.x, rr.y, tx), R1 = lerp_pair(rr.z, rr.w, tx);
                                    const float Rv = R0 + (R1 - R0) * pt.y;
                                    const float wgt = sm.w / L;
                                    contrib = Rv * lam * wgt;
                                }
                                const bool ok = L > kEpsilon;
                                sum[i] += ok ? contrib : 0.0f;
                                used[i] += ok ? 1 : 0;
                            }
                        }
                        float out[G];
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            const int at = zz * width + xw + i;
                            float value = band[at];
                            if (used[i] != 0) value += (kMask ? band[at + kMaskOff] * sum[i] : sum[i]) / (float)used[i];
                            out[i] = S::round(value);
                            band[at] = out[i];
                        }
                        if (q == nbp - 1)
                            IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw, out);
                    }
                } else if (kThreads % gpr == 0) {
                    // every thread keeps the same x group (and plane) in every band row
                    const int xw = (tid % gpr) * G, dz = kThreads / gpr;
                    int p = 0, x0 = xw;
                    if (np > 1) {
                        p = xw / nx;
                        x0 = xw - p * nx;
                    }
                    const float2 pt = pterm[p];
                    const float4* pr = recs + p * kRecStride * row_len;
                    float xt[G];
#pragma unroll
                    for (int i = 0; i < G; ++i) xt[i] = x_term(x0 + i, pt.x);
#pragma unroll 1
                    for (int zz = tid / gpr; zz < nrow; zz += dz) {
                        const float zt = z_term(zz);
                        float out[G];
                        float* cell = band + zz * width + xw;
                        const uint32_t vid = (uint32_t)((size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw);
                        if constexpr (G == 4 && kFast && !kSR && sizeof(typename S::T) == 2) {
                            // 16-bit storage: the four updates rounded in two packed
                            // conversions, the packed group stored as it is
                            const float4 in = *reinterpret_cast<const float4*>(cell);
                            const float4 mk = kMask ? *reinterpret_cast<const float4*>(cell + kMaskOff)
                                                    : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
                            const float4* prp = pr + pad_l;
                            out[0] = bp_update_fast<S, kMask, false, false>(prp, xt[0], zt, m12, in.x, mk.x, 0u, 0u);
                            out[1] = bp_update_fast<S, kMask, false, false>(prp, xt[1], zt, m12, in.y, mk.y, 0u, 0u);
                            out[2] = bp_update_fast<S, kMask, false, false>(prp, xt[2], zt, m12, in.z, mk.z, 0u, 0u);
                            out[3] = bp_update_fast<S, kMask, false, false>(prp, xt[3], zt, m12, in.w, mk.w, 0u, 0u);
                            const typename IO::V packed = IO::round4(out);
                            *reinterpret_cast<float4*>(cell) = make_float4(out[0], out[1], out[2], out[3]);
                            if (q == nbp - 1)
                                IO::store_packed(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw,
                                                 packed);
                            continue;
                        } else if constexpr (G == 4) {
                            // four voxels, 16-byte aligned (width and xw are multiples of 4)
                            const float4 in = *reinterpret_cast<const float4*>(cell);
                            const float4 mk = kMask ? *reinterpret_cast<const float4*>(cell + kMaskOff)
                                                    : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
                            out[0] = update(pr, pt.y, xt[0], zt, in.x, mk.x, vid);
                            out[1] = update(pr, pt.y, xt[1], zt, in.y, mk.y, vid + 1u);
                            out[2] = update(pr, pt.y, xt[2], zt, in.z, mk.z, vid + 2u);
                            out[3] = update(pr, pt.y, xt[3], zt, in.w, mk.w, vid + 3u);
                            *reinterpret_cast<float4*>(cell) = make_float4(out[0], out[1], out[2], out[3]);
                        } else {
#pragma unroll
                            for (int i = 0; i < G; ++i) {
                                out[i] = update(pr, pt.y, xt[i], zt, cell[i], kMask ? cell[i + kMaskOff] : 1.0f,
                                                vid + (uint32_t)i);
                                cell[i] = out[i];
                            }
                        }
                        if (q == nbp - 1)
                            IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw, out);
                    }
                } else {
#pragma unroll 1
                    for (int gi = tid; gi < ng; gi += kThreads) {
                        const int zz = gi / gpr, xw = (gi - zz * gpr) * G;
                        int p = 0, x0 = xw;
                        if (np > 1) {
                            p = xw / nx;
                            x0 = xw - p * nx;
                        }
                        const float2 pt = pterm[p];
                        const float4* pr = recs + p * kRecStride * row_len;
                        const float zt = z_term(zz);
                        float out[G];
                        const uint32_t vid = (uint32_t)((size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw);
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            const int c = zz * width + xw + i;
                            out[i] = update(pr, pt.y, x_term(x0 + i, pt.x), zt, band[c], kMask ? band[c + kMaskOff] : 1.0f,
                                            vid + (uint32_t)i);
                            band[c] = out[i];
                        }
                        if (q == nbp - 1)
                            IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw, out);
                    }
                }
                if (nbp > 1) __syncthreads();
            }
            if (nbp == 1) __syncthreads();
        }

        if (bi + 1 < nbands) prefetch(nb0, nb1, kSplit, kPre);

        // 3. forward projections of the current block through this band
        if (nfp > 0) {
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int id = k * kThreads + tid;
                if (id < nrays) plane_advance<kCap, kFast, kMask, kMaskOff>(ray[k], band, width, nx, nz, z0, z1);
            }
            if (!kCap && bi + 1 < nbands) {
                // re-base the LDS offsets on the next band's first row
                int n0, n1;
                band_range(bi + 1, n0, n1);
                const int shift = (z0 - n0) * width * 4;
#pragma unroll
                for (int k = 0; k < RPT; ++k) ray[k].li += shift;
            }
        }
        __syncthreads();
    }

    // ---- residual (+ long-object compensation) per pixel ----
    // oversampled: every traced sub-ray's sums go through LDS (free after the
    // last band); the pixel's first slot adds sub-ray (si, sj)'s sums, those of
    // (si, 0), for si, then sj, as the generic kernel does
    float2* part = reinterpret_cast<float2*>(lds);
    if (kOver) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int id = k * kThreads + tid;
            if (id < nrays) part[id] = make_float2(ray[k].acc, ray[k].accl);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int id = k * kThreads + tid;
        int q, p, w, u, sub;
        if (id < nrays && decode(id, q, p, w, u, sub) && sub == 0) {
            const ETN_GLOBAL FusedFp& f = pass->fp[q];
            float acc = 0.0f, accl = 0.0f;
            if (kOver) {
                for (int si = 0; si < fp_n; ++si) {
                    const float2 pa = part[id + si];
                    for (int sj = 0; sj < fp_n; ++sj) {
                        acc += pa.x;
                        accl += pa.y;
                    }
                }
            } else {
                acc += ray[k].acc;
                accl += ray[k].accl;
            }
            const float total = (float)(fp_n * fp_n);
            float pv = 0.0f + acc / total;
            float l = 0.0f + accl / total;
            if (f.use_loc) {
                pv = l > kEpsilon ? pv * f.loc_c / l : 0.0f;
                l = f.loc_c;
            }
            const size_t i = (size_t)w * nu + u;
            global_ptr(f.residual)[i] = f.raw ? pv : (kEarly ? meas[k] : measured_at(f, i)) - pv;
            global_ptr(f.length)[i] = l;
        }
    }

    // ---- rows whose rays miss the volume: zero line integral and length ----
    for (int p = 0; p < np; ++p)
        for (int d = j + p; d < fv.ndead; d += fv.ny) {
            const int w = global_ptr(fv.dead_rows)[d];
            for (int q = 0; q < nfp; ++q) {
                const ETN_GLOBAL FusedFp& f = pass->fp[q];
                for (int u = tid; u < nu; u += kThreads) {
                    float pv = 0.0f + 0.0f / 1.0f;
                    float l = 0.0f + 0.0f / 1.0f;
                    if (f.use_loc) {
                        pv = l > kEpsilon ? pv * f.loc_c / l : 0.0f;
                        l = f.loc_c;
                    }
                    const size_t i = (size_t)w * nu + u;
                    global_ptr(f.residual)[i] = f.raw ? pv : measured_at(f, i) - pv;
                    global_ptr(f.length)[i] = l;
                }
            }
        }
}

template <typename S, bool kVec, bool kCap, bool kFast, bool kMask, bool kOver, int kThreads, bool kSR = false,
          int kRes = 0>
void launch_fused_shape(const FusedLaunch& l, hipStream_t stream)
{
    using Shape = typename FusedShapeOf<kThreads, (int)sizeof(typename S::T), kRes>::type;
    if (l.lds > sizeof(float) * (size_t)Shape::kLds) fail("fused plane kernel: tile exceeds the LDS array");
    using Kernel = void (*)(const FusedVolume*, const FusedWork*, const FusedPass* const*, int);
    Kernel kernel;
    switch (l.rpt) {
        case 1: kernel = fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 1, kSR, kRes>; break;
        case 2: kernel = fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 2, kSR, kRes>; break;
        case 4: kernel = fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 4, kSR, kRes>; break;
        default: kernel = fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 8, kSR, kRes>; break;
    }
    if (l.resident) {
        int per_cu = 0, dev = 0, cus = 0;
        ETN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0));
        ETN_HIP(hipGetDevice(&dev));
        ETN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        *l.resident = per_cu * cus;
        return;
    }
    kernel<<<dim3(l.groups), dim3(kThreads), 0, stream>>>(l.vols, l.work, l.passes, l.band_rows);
}

// The oversampled kernels exist for 512-thread workgroups only.
template <typename S, bool kVec, bool kCap, bool kFast, bool kMask, bool kOver>
void launch_fused_rpt(const FusedLaunch& l, hipStream_t stream)
{
    if (l.threads == 512) {
        launch_fused_shape<S, kVec, kCap, kFast, kMask, kOver, 512>(l, stream);
        return;
    }
    if constexpr (!kOver) {
        if (l.threads == 256) {
            launch_fused_shape<S, kVec, kCap, kFast, kMask, false, 256>(l, stream);
            return;
        }
    }
    fail("fused plane kernel: workgroups have 512 threads, or 256 without oversampling");
}

template <typename S, bool kFast, bool kMask, bool kOver>
void launch_fused_typed(bool vec, bool cap, const FusedLaunch& l, hipStream_t stream)
{
    if (l.res == 3) {
        if constexpr (!kOver) {
            if (vec && !cap && !l.sr && l.threads == 512 && l.band_rows > 0) {
                launch_fused_shape<S, true, false, kFast, kMask, false, 512, false, 3>(l, stream);
                return;
            }
        }
        fail("fused plane kernel: the 160 KiB LDS back projection needs nx a multiple of 4, no oversampling, "
             "stochastic rounding or step limit, and 512 threads");
    }
    if (l.res != 0) {
        if constexpr (kFast && !kOver) {
            if (vec && !cap && !l.sr) {
                if (l.res == 2 && l.threads == 512) {
                    launch_fused_shape<S, true, false, true, kMask, false, 512, false, 2>(l, stream);
                    return;
                }
                if (l.res == 4 && l.threads == 512 && l.band_rows > 0) {
                    launch_fused_shape<S, true, false, true, kMask, false, 512, false, 4>(l, stream);
                    return;
                }
                if (l.res == 1 && l.threads == 512) {
                    launch_fused_shape<S, true, false, true, kMask, false, 512, false, 1>(l, stream);
                    return;
                }
                if (l.res == 1 && l.threads == 256) {
                    launch_fused_shape<S, true, false, true, kMask, false, 256, false, 1>(l, stream);
                    return;
                }
            }
        }
        fail("fused plane kernel: the register back projection needs fast arithmetic, nx a multiple of 4, no "
             "oversampling, stochastic rounding or step limit, and 512 threads for the resident record sets");
    }
    if (l.sr) {
        // stochastic rounding: 16-bit storage on the fast kernels without a
        // mask, oversampling or the step limit, nx a multiple of 4
        if constexpr (kFast && !kMask && !kOver && !std::is_same<S, F32>::value) {
            if (vec && !cap) {
                if (l.threads == 512) launch_fused_shape<S, true, false, true, false, false, 512, true>(l, stream);
                else launch_fused_shape<S, true, false, true, false, false, 256, true>(l, stream);
                return;
            }
        }
        fail("stochastic rounding needs 16-bit storage, fast arithmetic, no mask, no oversampling, nx a multiple "
             "of 4 and rays that cannot reach the 2048-step limit");
    }
    if (vec && !cap) launch_fused_rpt<S, true, false, kFast, kMask, kOver>(l, stream);
    else if (vec) launch_fused_rpt<S, true, true, kFast, kMask, kOver>(l, stream);
    else if (!cap) launch_fused_rpt<S, false, false, kFast, kMask, kOver>(l, stream);
    else launch_fused_rpt<S, false, true, kFast, kMask, kOver>(l, stream);
}

}  // namespace

// The kernels of one arithmetic mode and storage type, instantiated in a
// translation unit of their own (fused_inst.hip, built once per pair) so that
// the variants compile in parallel.  mask: prior-knowledge masks; over:
// oversampled FP and BP; vec: every nx is a multiple of 4; cap: the 2048-step
// limit can bind.
void launch_fused_fast_F32(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream);
void launch_fused_fast_F16(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream);
void launch_fused_fast_BF16(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream);
void launch_fused_exact_F32(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream);
void launch_fused_exact_F16(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream);
void launch_fused_exact_BF16(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream);

}  // namespace dev
}  // namespace etn
