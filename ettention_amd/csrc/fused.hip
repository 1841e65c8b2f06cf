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
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "block_iterative.hpp"
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

// Workgroup shapes.  512 threads with a 64 KiB LDS array run two workgroups
// per CU; 256 threads with 40 KiB run four, whose independent barrier phases
// overlap better when a pass carries many projections, at half the band.
// kBand: band voxels (the vector path prefetches 6 groups of 4 per thread).
// With a prior-knowledge mask the band holds kBand / 2 voxels and the voxels'
// mask weights follow at LDS offset kBand / 2 (floats).
template <int kThreads>
struct FusedShape;
template <>
struct FusedShape<512> {
    static constexpr int kLds = 16384;
    static constexpr int kBand = 12288;
};
template <>
struct FusedShape<256> {
    static constexpr int kLds = 10240;
    static constexpr int kBand = 6144;
};
constexpr int kMaxThreads = 512;
__host__ __device__ inline int shape_lds_floats(int threads) { return threads == 256 ? FusedShape<256>::kLds : FusedShape<512>::kLds; }
__host__ __device__ inline int shape_band_floats(int threads, bool mask)
{
    const int b = threads == 256 ? FusedShape<256>::kBand : FusedShape<512>::kBand;
    return mask ? b / 2 : b;
}

struct FusedFp {
    float src[3], h[3], v[3], ray[3];
    const float* measured;       // nv x nu
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
// at offset shape_band_floats(threads, true) the band's mask weights), then
// (16-byte aligned) the staged detector records of each plane and row set
// (row_floats per row set: 8 floats per padded column with exact arithmetic,
// 4 with fast; one row set per BP pattern row) and a (m4 py, ty) pair per
// plane and row set.
__host__ __device__ inline int fused_lds_floats(int planes, int band, int nx, int row_floats, bool mask, int threads,
                                                int rowsets)
{
    return (mask ? shape_band_floats(threads, true) : 0) + ((band * planes * nx + 3) & ~3) +
           planes * rowsets * (row_floats + 2);
}

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
// dropped when nx + nz <= 2049, since every step moves vx or vz monotonically by
// one, so a ray can take at most nx + nz - 1 steps inside the volume.
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
// access granularity of the band transfers.
template <typename S, bool kVec>
struct GroupIo;
template <>
struct GroupIo<F32, true> {
    static constexpr int G = 4;
    using V = float4;
    static __device__ __forceinline__ V load(const ETN_GLOBAL float* p, size_t i)
    {
        return *reinterpret_cast<const ETN_GLOBAL float4*>(p + i);
    }
    static __device__ __forceinline__ void unpack(const V& v, float* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
    static __device__ __forceinline__ void store(ETN_GLOBAL float* p, size_t i, const float* f)
    {
        *reinterpret_cast<ETN_GLOBAL float4*>(p + i) = make_float4(f[0], f[1], f[2], f[3]);
    }
};
template <typename S16>
struct GroupIo16 {   // 4 x 16-bit voxels as one 8-byte access
    static constexpr int G = 4;
    using V = uint2;
    static __device__ __forceinline__ V load(const ETN_GLOBAL typename S16::T* p, size_t i)
    {
        return *reinterpret_cast<const ETN_GLOBAL uint2*>(p + i);
    }
    static __device__ __forceinline__ void unpack(const V& v, float* o)
    {
        o[0] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 0);
        o[1] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 1);
        o[2] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 2);
        o[3] = S16::load(reinterpret_cast<const typename S16::T*>(&v), 3);
    }
    static __device__ __forceinline__ void store(ETN_GLOBAL typename S16::T* p, size_t i, const float* f)
    {
        V v;
        typename S16::T* t = reinterpret_cast<typename S16::T*>(&v);
        for (int k = 0; k < 4; ++k) S16::store(t, k, f[k]);
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
template <typename S, bool kVec, bool kCap, bool kFast, bool kMask, bool kOver, int kThreads, int RPT>
__global__ void __launch_bounds__(kThreads, RPT <= 2 ? 4 : 1)
    fused_plane_kernel(const FusedVolume* __restrict__ vols, const FusedWork* __restrict__ work,
                       const FusedPass* const* __restrict__ passes)
{
    using IO = GroupIo<S, kVec>;
    using Shape = FusedShape<kThreads>;
    constexpr int G = IO::G;
    constexpr int kMaskOff = Shape::kBand / 2;          // LDS floats from a voxel to its mask weight
    constexpr int kBandCap = kMask ? kMaskOff : Shape::kBand;
    constexpr int kGroups = kBandCap / (G * kThreads);  // band groups per thread (upper bound)
    // a fixed-size array puts the band at LDS address 0 (the host sizes every
    // tile to fit it; the VGPR budget limits the workgroups per CU as much)
    __shared__ __attribute__((aligned(16))) float lds[Shape::kLds];
    const FusedWork item = work[blockIdx.x];
    const ETN_GLOBAL FusedVolume& fv = global_ptr(vols)[item.member];
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
    float4* recs = reinterpret_cast<float4*>(lds + (kMask ? kMaskOff : 0) + ((fv.band * width + 3) & ~3));
    constexpr int kRecStride = kFast ? 1 : 2;   // float4 per column record
    float2* pterm = reinterpret_cast<float2*>(recs + np * rowsets * kRecStride * row_len);
    ETN_GLOBAL typename S::T* vol = global_ptr(static_cast<typename S::T*>(fv.data));
    const ETN_GLOBAL uint8_t* mask = kMask ? global_ptr(fv.mask) : nullptr;
    // MaskVolume weight of a byte, as the generic kernels compute it
    auto weight = [](uint32_t m) { return (float)m / 255.0f; };
    const size_t plane_stride = (size_t)nx * fv.ny;
    const size_t plane_base = (size_t)j * nx;
    const int gpr = width / G;                          // groups per band row

    const int nfp = pass->nfp, nbp = pass->nbp;
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

    // ---- ray set-up: dda_begin exactly as the generic FP kernel ----
    PlaneRay ray[RPT];
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

    const int nbands = (nz + fv.band - 1) / fv.band;

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
    auto stage_rows = [&](int q) {
        const ETN_GLOBAL FusedBp& bp = pass->bp[q];
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
                // R(tx) = A + B tx with A, B the row pair interpolated at ty
                recs[e] = make_float4(fmaf(v[2] - v[0], ty, v[0]), fmaf(v[3] - v[1], ty, v[1]),
                                      fmaf(v[6] - v[4], ty, v[4]), fmaf(v[7] - v[5], ty, v[5]));
            } else {
                recs[2 * e] = make_float4(v[0], v[1], v[2], v[3]);
                recs[2 * e + 1] = make_float4(v[4], v[5], v[6], v[7]);
            }
        }
        if (tid < np * rowsets) {
            const int p = tid / rowsets, ys = tid - p * rowsets;
            const float dy = kOver ? bp.samples[ys * bp.per_dir].y : bp.sample.y;
            float py, ty;
            plane_v(bp, p, dy, py, ty);
            pterm[tid] = make_float2(bp.m[4] * (py + dy), ty);
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
    uint32_t mpre[kMask ? kPre : 1];                    // the groups' 4 mask bytes
    auto prefetch = [&](int z0, int z1) {
        if (!kVec) return;
        const int ng = (z1 - z0) * gpr;
#pragma unroll
        for (int k = 0; k < kPre; ++k) {
            const int gi = k * kThreads + tid;
            if (gi < ng) {
                const int zz = gi / gpr, xg = gi - zz * gpr;
                const size_t at = (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xg * G;
                pre[k] = IO::load(vol, at);
                if (kMask) mpre[k] = mask ? *reinterpret_cast<const ETN_GLOBAL uint32_t*>(mask + at) : 0xffffffffu;
            }
        }
    };

    int z0, z1;
    band_range(0, z0, z1);
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
    prefetch(z0, z1);
    if (nbp == 1) stage_rows(0);     // made visible by the first band's barrier

    for (int bi = 0; bi < nbands; ++bi) {
        band_range(bi, z0, z1);
        const int ng = (z1 - z0) * gpr;
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
        // workgroup stores later)
        if (bi + 1 < nbands) {
            int n0, n1;
            band_range(bi + 1, n0, n1);
            prefetch(n0, n1);
        }
        __syncthreads();

        // 2. back projections of the previous block, in block order; each thread
        //    updates (and finally stores) the groups it staged
        for (int q = 0; q < nbp; ++q) {
            if (nbp > 1) {
                stage_rows(q);
                __syncthreads();
            }
            // the pass lives in global memory: copy what the loops use into registers
            const ETN_GLOBAL FusedBp& bp = pass->bp[q];
            const float m0 = bp.m[0], m8 = bp.m[8], m12 = bp.m[12], lam = bp.lambda;
            const float4 smp = bp.sample;
            const float chord = smp.w, lc = lam * chord;
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
            auto update = [&](const float4* pr, float ty, float xt, float zt, float value, float mv) {
                if (kFast) {
                    // u in the reference's order (u is a few hundred pixels, so a
                    // reordered sum would move the weights by ~1e-4); fract without
                    // the clamp below 1 (at tx == 1 the interpolation gives the next
                    // column's value, its limit).  The staged rows are padded (on the
                    // host) beyond every column the member's projections reach, so the
                    // index needs no clamp.
                    float u = xt + zt;
                    u = u + m12 * 1.0f;
                    u = u - 0.5f;
                    const float bu = floorf(u);
                    const float tx = u - bu;
                    const int i = (int)bu + pad_l;
                    const float4 c = pr[i];
                    const float Rv = fmaf(c.y, tx, c.x), L = fmaf(c.w, tx, c.z);
                    const float contrib = Rv * lc * __builtin_amdgcn_rcpf(L);
                    const float next = kMask ? fmaf(mv, contrib, value) : value + contrib;
                    value = (L > kEpsilon) ? next : value;
                    return S::round(value);
                }
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
                    float out[G];
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const float cx = s.bmin[0] + ((float)(x0 + i) + 0.5f) * s.vs[0];
                        float sum = 0.0f;
                        int used = 0;
#pragma unroll 1
                        for (int t = 0; t < nsmp; ++t) {
                            const float4 sm = bp.samples[t];
                            const int ys = t / pd;
                            const float2 pt = pterm[p * rowsets + ys];
                            float u = m0 * (cx + sm.x) + pt.x;
                            u = u + m8 * (cz + sm.z);
                            u = u + m12 * 1.0f;
                            float L, contrib;
                            if (kFast) {
                                u = u - 0.5f;
                                const float bu = floorf(u);
                                const float tx = u - bu;
                                const float4 c = recs[(p * rowsets + ys) * row_len + (int)bu + pad_l];
                                const float Rv = fmaf(c.y, tx, c.x);
                                L = fmaf(c.w, tx, c.z);
                                contrib = Rv * (lam * sm.w) * __builtin_amdgcn_rcpf(L);
                            } else {
                                // bp_voxel_sum's bilinear lookups and weight, operation for operation
                                float bu;
                                const float tx = cl_fract(u - 0.5f, &bu);
                                const int i = padded_index((int)bu, nu);
                                const float4* pr = recs + (p * rowsets + ys) * 2 * row_len;
                                const float4 rr = pr[2 * i], ll = pr[2 * i + 1];
                                const float L0 = lerp_pair(ll.x, ll.y, tx), L1 = lerp_pair(ll.z, ll.w, tx);
                                L = L0 + (L1 - L0) * pt.y;
                                const float R0 = lerp_pair(rr.x, rr.y, tx), R1 = lerp_pair(rr.z, rr.w, tx);
                                const float Rv = R0 + (R1 - R0) * pt.y;
                                const float wgt = sm.w / L;
                                contrib = Rv * lam * wgt;
                            }
                            const bool ok = L > kEpsilon;
                            sum += ok ? contrib : 0.0f;
                            used += ok ? 1 : 0;
                        }
                        const int at = zz * width + xw + i;
                        float value = band[at];
                        if (used != 0) value += (kMask ? band[at + kMaskOff] * sum : sum) / (float)used;
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
                    if constexpr (G == 4) {
                        // four voxels, 16-byte aligned (width and xw are multiples of 4)
                        const float4 in = *reinterpret_cast<const float4*>(cell);
                        const float4 mk = kMask ? *reinterpret_cast<const float4*>(cell + kMaskOff)
                                                : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
                        out[0] = update(pr, pt.y, xt[0], zt, in.x, mk.x);
                        out[1] = update(pr, pt.y, xt[1], zt, in.y, mk.y);
                        out[2] = update(pr, pt.y, xt[2], zt, in.z, mk.z);
                        out[3] = update(pr, pt.y, xt[3], zt, in.w, mk.w);
                        *reinterpret_cast<float4*>(cell) = make_float4(out[0], out[1], out[2], out[3]);
                    } else {
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            out[i] = update(pr, pt.y, xt[i], zt, cell[i], kMask ? cell[i + kMaskOff] : 1.0f);
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
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const int c = zz * width + xw + i;
                        out[i] = update(pr, pt.y, x_term(x0 + i, pt.x), zt, band[c], kMask ? band[c + kMaskOff] : 1.0f);
                        band[c] = out[i];
                    }
                    if (q == nbp - 1)
                        IO::store(vol, (size_t)(z0 + zz) * plane_stride + plane_base + (size_t)xw, out);
                }
            }
            if (nbp > 1) __syncthreads();
        }
        if (nbp == 1) __syncthreads();

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
            global_ptr(f.residual)[i] = f.raw ? pv : global_ptr(f.measured)[i] - pv;
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
                    global_ptr(f.residual)[i] = f.raw ? pv : global_ptr(f.measured)[i] - pv;
                    global_ptr(f.length)[i] = l;
                }
            }
        }
}

struct FusedLaunch {
    int threads;                 // workgroup shape: 512 or 256
    int rpt;
    int groups;
    const FusedVolume* vols;
    const FusedWork* work;
    const FusedPass* const* passes;
    size_t lds;
};

template <typename S, bool kVec, bool kCap, bool kFast, bool kMask, bool kOver, int kThreads>
void launch_fused_shape(const FusedLaunch& l, hipStream_t stream)
{
    if (l.lds > sizeof(float) * (size_t)FusedShape<kThreads>::kLds) fail("fused plane kernel: tile exceeds the LDS array");
    const dim3 grid(l.groups), block(kThreads);
    switch (l.rpt) {
        case 1: fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 1><<<grid, block, 0, stream>>>(l.vols, l.work, l.passes); break;
        case 2: fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 2><<<grid, block, 0, stream>>>(l.vols, l.work, l.passes); break;
        case 4: fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 4><<<grid, block, 0, stream>>>(l.vols, l.work, l.passes); break;
        default: fused_plane_kernel<S, kVec, kCap, kFast, kMask, kOver, kThreads, 8><<<grid, block, 0, stream>>>(l.vols, l.work, l.passes); break;
    }
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
    if (vec && !cap) launch_fused_rpt<S, true, false, kFast, kMask, kOver>(l, stream);
    else if (vec) launch_fused_rpt<S, true, true, kFast, kMask, kOver>(l, stream);
    else if (!cap) launch_fused_rpt<S, false, false, kFast, kMask, kOver>(l, stream);
    else launch_fused_rpt<S, false, true, kFast, kMask, kOver>(l, stream);
}

template <bool kMask, bool kOver>
void launch_fused_fast(etn_voxel_type vt, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream)
{
    if (vt == ETN_VOXEL_FLOAT32) launch_fused_typed<F32, true, kMask, kOver>(vec, cap, l, stream);
    else if (vt == ETN_VOXEL_FLOAT16) launch_fused_typed<F16, true, kMask, kOver>(vec, cap, l, stream);
    else launch_fused_typed<BF16, true, kMask, kOver>(vec, cap, l, stream);
}

}  // namespace

// vec: every member's nx is a multiple of 4; cap: some member has nx + nz > 2049;
// exact: the generic kernels' operations and rounding order; mask: the launch
// carries prior-knowledge masks; over: oversampled FP and BP
void launch_fused(etn_voxel_type vt, bool vec, bool cap, bool exact, bool mask, bool over, const FusedLaunch& l,
                  hipStream_t stream)
{
    if (exact && over) {
        if (mask) {
            if (vt == ETN_VOXEL_FLOAT32) launch_fused_typed<F32, false, true, true>(vec, cap, l, stream);
            else if (vt == ETN_VOXEL_FLOAT16) launch_fused_typed<F16, false, true, true>(vec, cap, l, stream);
            else launch_fused_typed<BF16, false, true, true>(vec, cap, l, stream);
        } else {
            if (vt == ETN_VOXEL_FLOAT32) launch_fused_typed<F32, false, false, true>(vec, cap, l, stream);
            else if (vt == ETN_VOXEL_FLOAT16) launch_fused_typed<F16, false, false, true>(vec, cap, l, stream);
            else launch_fused_typed<BF16, false, false, true>(vec, cap, l, stream);
        }
    } else if (exact) {
        if (mask) {
            if (vt == ETN_VOXEL_FLOAT32) launch_fused_typed<F32, false, true, false>(vec, cap, l, stream);
            else if (vt == ETN_VOXEL_FLOAT16) launch_fused_typed<F16, false, true, false>(vec, cap, l, stream);
            else launch_fused_typed<BF16, false, true, false>(vec, cap, l, stream);
        } else {
            if (vt == ETN_VOXEL_FLOAT32) launch_fused_typed<F32, false, false, false>(vec, cap, l, stream);
            else if (vt == ETN_VOXEL_FLOAT16) launch_fused_typed<F16, false, false, false>(vec, cap, l, stream);
            else launch_fused_typed<BF16, false, false, false>(vec, cap, l, stream);
        }
    } else if (over) {
        if (mask) launch_fused_fast<true, true>(vt, vec, cap, l, stream);
        else launch_fused_fast<false, true>(vt, vec, cap, l, stream);
    } else {
        if (mask) launch_fused_fast<true, false>(vt, vec, cap, l, stream);
        else launch_fused_fast<false, false>(vt, vec, cap, l, stream);
    }
    hip_check(hipGetLastError(), "fused plane kernel launch");
}

}  // namespace dev

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct Tile {
    int planes = 1;
    int band = 1;
};

// One reconstruction driven by the engine: its plane/row tables, its
// residual images and the passes of one iteration under its current tile.
struct MemberPlan {
    BlockIterativeReconstructionOperator* op = nullptr;
    int nx = 0, ny = 0, nz = 0, nu = 0, nv = 0;
    size_t pixels = 0;
    float lambda = 0.0f;
    bool use_loc = false;
    int rows_per_plane = 0;
    int row_len = 0, pad_l = 0;        // staged detector columns, zero columns left of column 0
    int row_floats = 0;                // LDS floats of one plane's staged rows
    const uint8_t* mask = nullptr;     // prior-knowledge mask (device), or null
    int threads = dev::kMaxThreads;    // workgroup shape of the engine's launches
    int fp_n = 1, fp_shift = 0;        // forward sub-rays per pixel axis; log2 of ray slots (columns si) per pixel
    int bp_samples = 1, rowsets = 1;   // back projection samples per voxel; pattern rows
    bool oversampled() const { return fp_n != 1 || bp_samples != 1; }
    int ndead = 0;
    uint32_t bmax = 0;                 // largest block
    DeviceBuffer rows;                 // row_lo[ny] | row_hi[ny] | dead rows
    DeviceBuffer images;               // residual + length per block slot, two parities
    Tile tile;
    std::vector<dev::FusedPass> passes;
    int offset = 0;                    // launch that carries the member's first pass
};

int select_rpt(int rays, int threads)
{
    int rpt = 1;
    while (rpt * threads < rays && rpt < dev::kMaxRaysPerThread) rpt *= 2;
    return rpt;
}

// Largest band for `planes` planes per workgroup of `threads`, 0 when the tile does not fit.
int max_band(const MemberPlan& m, int planes, int threads)
{
    if (planes < 1 || planes > dev::kMaxPlanes || (planes & (planes - 1)) || m.ny % planes) return 0;
    if ((planes * m.rows_per_plane * m.nu << m.fp_shift) > threads * dev::kMaxRaysPerThread) return 0;
    if (m.oversampled() && threads != dev::kMaxThreads) return 0;   // oversampled kernels: 512 threads only
    const bool mask = m.mask != nullptr;
    const int masked = mask ? dev::shape_band_floats(threads, true) : 0;
    int band = std::min(m.nz, (dev::shape_lds_floats(threads) - masked - 3 -
                               planes * m.rowsets * (m.row_floats + 2)) / (planes * m.nx));
    band = std::min(band, dev::shape_band_floats(threads, mask) / (planes * m.nx));
    return std::max(band, 0);
}

int max_band(const MemberPlan& m, int planes) { return max_band(m, planes, m.threads); }

bool tile_fits(const MemberPlan& m, const Tile& t)
{
    return t.band >= 1 && t.band <= max_band(m, t.planes) &&
           dev::fused_lds_floats(t.planes, t.band, m.nx, m.row_floats, m.mask != nullptr, m.threads, m.rowsets) <=
               dev::shape_lds_floats(m.threads);
}

// Without timing: the fewest planes whose rays fill a workgroup, the tallest band.
Tile default_tile(const MemberPlan& m)
{
    Tile best{1, max_band(m, 1)};
    for (int p = 1; p <= dev::kMaxPlanes; p *= 2) {
        const int b = max_band(m, p);
        if (b < 1) continue;
        best = {p, b};
        if ((p * m.rows_per_plane * m.nu << m.fp_shift) >= m.threads) break;
    }
    return best;
}

std::vector<Tile> candidate_tiles(const MemberPlan& m)
{
    std::vector<Tile> out;
    for (int p = 1; p <= dev::kMaxPlanes; p *= 2) {
        const int b = max_band(m, p);
        if (b < 1) continue;
        out.push_back({p, b});
        if (b >= 2) out.push_back({p, (b + 1) / 2});
    }
    return out;
}

// The fused kernels sweep the whole volume.  A z-slab split
// (hardware.subVolumeCount) restarts the reference's forward projection at
// every slab face and sums the slab parts; sweeping the volume in one pass
// takes the same steps and only regroups those sums and the voxel-centre
// arithmetic, so it is allowed with fast arithmetic and left to the generic
// kernels with exact arithmetic.
VolumeProperties whole_volume(const VolumeProperties& vp)
{
    etn_volume_geometry g = vp.desc();
    g.subvolume_count = 1;
    return VolumeProperties(g);
}

// The fast lookup skips the column clamp: stage every column u - 0.5 can
// floor to for a voxel centre (and every sample of the BP pattern) under any
// of the geometries, plus one column either side.  False when the span is
// unbounded.
bool fast_row_span(const VolumeProperties& vp, const std::vector<const ParallelScannerGeometry*>& geoms, int nu,
                   int bp_s, int& pad_l, int& row_len)
{
    const Vec3f vs = vp.voxel_size(), lo = vp.bbox_min(), hi = vp.bbox_max();
    double umin = 1e300, umax = -1e300;
    for (const ParallelScannerGeometry* g : geoms) {
        const Matrix4x4 pm = g->projection_matrix();
        const float* mv = pm.data();
        for (const VoxelSample& smp : bp_sample_pattern(*g, vs, (uint32_t)bp_s))
            for (int c = 0; c < 8; ++c) {
                const double px = ((c & 1) ? hi.x - 0.5 * vs.x : lo.x + 0.5 * vs.x) + smp.dx;
                const double py = ((c & 2) ? hi.y - 0.5 * vs.y : lo.y + 0.5 * vs.y) + smp.dy;
                const double pz = ((c & 4) ? hi.z - 0.5 * vs.z : lo.z + 0.5 * vs.z) + smp.dz;
                const double u = mv[0] * px + mv[4] * py + mv[8] * pz + mv[12];
                umin = std::min(umin, u);
                umax = std::max(umax, u);
            }
    }
    if (!(umin > -1e6 && umax < 1e6)) return false;
    const int bu_lo = (int)std::floor(umin - 0.5) - 1, bu_hi = (int)std::floor(umax - 0.5) + 1;
    pad_l = std::max(dev::kRowPad, -bu_lo);
    row_len = pad_l + std::max(nu + dev::kRowPad, bu_hi + 1);
    return true;
}

// Eligibility of one reconstruction and its plane/row tables; false when it
// has no fused form.
bool plan_member(BlockIterativeReconstructionOperator& op, const etn_reconstruction_desc& d, bool exact,
                 MemberPlan& out)
{
    if (exact && op.getReconstructedVolume()->properties().subvolume_count() != 1) return false;
    const VolumeProperties whole = whole_volume(op.getReconstructedVolume()->properties());
    const VolumeProperties& vp = whole;
    // oversampling: up to 4 x 4 forward sub-rays and a 4 x 4 back projection pattern
    const int fp_n = (int)(d.fp_samples ? d.fp_samples : 1);
    const int bp_s = (int)(d.bp_samples ? d.bp_samples : 1);
    const int bp_pd = (int)std::lround(std::sqrt((double)bp_s));
    if (fp_n > dev::kMaxFusedFpPerAxis || bp_s > dev::kMaxFusedBpSamples || bp_pd * bp_pd != bp_s) return false;
    // forward sub-rays restarted at every slab face (the reference's slab-wise
    // FP) differ from one sweep by up to 2.4e-4 relative L2 after two SART
    // iterations (4 x 4 sub-rays, 2 or 4 slabs): beyond the FP32 bound, so an
    // oversampled FP over slabs stays on the generic kernels
    if (fp_n != 1 && op.getReconstructedVolume()->properties().subvolume_count() != 1) return false;
    int fp_shift = 0;
    while ((1 << fp_shift) < fp_n) ++fp_shift;     // ray slots per pixel: sub-ray columns si
    if (reinterpret_cast<uintptr_t>(d.mask) % 4) return false;   // masks are read 4 bytes at a time
    const int nx = (int)vp.nx(), ny = (int)vp.ny(), nz = (int)vp.nz();
    const int nu = (int)d.base_geometry.resolution[0], nv = (int)d.base_geometry.resolution[1];

    // every projection: rays in y planes, constant detector row per plane
    std::vector<int> plane_of_row;
    for (uint32_t b = 0; b < op.GetBlockCount(); ++b)
        for (uint32_t p = 0; p < op.getBlockSize(b); ++p) {
            const ParallelScannerGeometry& g = op.geometry(op.getPhysicalProjectionIndex(b, p));
            const Vec3f r = g.ray_direction();
            if (r.y != 0.0f || g.horizontal_pitch().y != 0.0f) return false;
            // oversampled FP: sub-rays of one column differ only in y
            if (fp_n > 1 && (g.vertical_pitch().x != 0.0f || g.vertical_pitch().z != 0.0f)) return false;
            // nonzero x/z components must have finite reciprocals (the plane
            // DDA relies on every res value staying at or below FLT_MAX)
            for (float c : {r.x, r.z})
                if (c != 0.0f && std::fabs(c) < 1e-30f) return false;
            const Matrix4x4 m = g.projection_matrix();
            const float* mv = m.data();
            if (mv[1] != 0.0f || mv[9] != 0.0f || mv[3] != 0.0f || mv[7] != 0.0f || mv[11] != 0.0f || mv[15] != 1.0f)
                return false;
            // plane of each detector row, with the FP's own arithmetic for every
            // sub-ray row sj (all must agree); -1 for rows whose sources lie
            // outside the box in y by a margin that makes every ray's y slab
            // interval empty (t_out <= -2^-20 * FLT_MAX)
            std::vector<int> map(nv);
            const SubVolume& sub = vp.subvolume(0);
            const float ivs = 1.0f / vp.voxel_size().y;
            const float margin = 1.0f / 1048576.0f;
            const float sd = fp_n == 1 ? 1.0f : 1.0f / (float)fp_n;
            const float off = fp_n == 1 ? 0.5f : sd / 2.0f;
            for (int w = 0; w < nv; ++w) {
                for (int sj = 0; sj < fp_n; ++sj) {
                    const float fx = (float)0 + off + sd * 0.0f;
                    const float fy = (float)w + off + sd * (float)sj;
                    const float ru = g.horizontal_pitch().y * fx, rv = g.vertical_pitch().y * fy;
                    float y = g.source_base().y + ru + rv;
                    int jj;
                    if (sub.bmin.y - y >= margin || y - sub.bmax.y >= margin) {
                        jj = -1;
                    } else {
                        if (y < sub.bmin.y || y > sub.bmax.y) return false;   // too close to call: generic kernels
                        y = std::fmin(std::fmax(y, sub.bmin.y), sub.bmax.y);
                        jj = (int)((y - sub.bmin.y) * ivs);
                        jj = std::max(jj, 0);
                        jj = std::min(jj, ny - 1);
                    }
                    if (sj == 0) map[w] = jj;
                    else if (map[w] != jj) return false;   // the pixel's sub-rays cross planes
                }
            }
            if (plane_of_row.empty()) plane_of_row = map;
            else if (plane_of_row != map) return false;
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
            if (seen[jj]) return false;        // rows of a plane must be contiguous
            seen[jj] = 1;
            lo[jj] = w;
            prev = jj;
        }
        hi[jj] = w + 1;
    }
    int rows_per_plane = 0;
    for (int jj = 0; jj < ny; ++jj) rows_per_plane = std::max(rows_per_plane, hi[jj] - lo[jj]);
    if (rows_per_plane < 1 || (rows_per_plane * nu << fp_shift) > dev::kMaxThreads * dev::kMaxRaysPerThread)
        return false;

    out.op = &op;
    out.nx = nx;
    out.ny = ny;
    out.nz = nz;
    out.nu = nu;
    out.nv = nv;
    out.pixels = (size_t)nu * nv;
    out.lambda = d.lambda;
    out.use_loc = d.use_long_object_compensation != 0;
    out.rows_per_plane = rows_per_plane;
    out.pad_l = dev::kRowPad;
    out.row_len = nu + 2 * dev::kRowPad;
    if (!exact) {
        std::vector<const ParallelScannerGeometry*> geoms;
        for (uint32_t b = 0; b < op.GetBlockCount(); ++b)
            for (uint32_t p = 0; p < op.getBlockSize(b); ++p)
                geoms.push_back(&op.geometry(op.getPhysicalProjectionIndex(b, p)));
        if (!fast_row_span(vp, geoms, nu, bp_s, out.pad_l, out.row_len)) return false;
    }
    out.row_floats = out.row_len * (exact ? 8 : 4);
    out.mask = d.mask;
    out.threads = dev::kMaxThreads;
    out.fp_n = fp_n;
    out.fp_shift = fp_shift;
    out.bp_samples = bp_s;
    out.rowsets = bp_pd;
    if (dev::fused_lds_floats(1, 1, nx, out.row_floats, out.mask != nullptr, out.threads, out.rowsets) >
        dev::shape_lds_floats(out.threads))
        return false;   // a z row and the rows fit
    if (max_band(out, 1) < 1) return false;   // not even one z row of one plane fits
    out.ndead = (int)dead.size();
    std::vector<int> table(lo);
    table.insert(table.end(), hi.begin(), hi.end());
    table.insert(table.end(), dead.begin(), dead.end());
    out.rows.allocate(sizeof(int) * table.size());
    ETN_HIP(hipMemcpy(out.rows.get(), table.data(), sizeof(int) * table.size(), hipMemcpyHostToDevice));
    out.bmax = 0;
    for (uint32_t b = 0; b < op.GetBlockCount(); ++b) out.bmax = std::max(out.bmax, op.getBlockSize(b));
    out.images.allocate(sizeof(float) * out.pixels * 2 * 2 * out.bmax);
    out.tile = default_tile(out);
    return true;
}

// The passes of one iteration: pass i back-projects what is left of the
// previous block and forward-projects a group of the current block.
void build_passes(MemberPlan& mb)
{
    BlockIterativeReconstructionOperator& op = *mb.op;
    mb.passes.clear();
    const int max_fp = std::max(1, std::min(dev::kMaxFusedFp, mb.threads * dev::kMaxRaysPerThread /
                                                                  ((mb.tile.planes * mb.rows_per_plane * mb.nu)
                                                                   << mb.fp_shift)));
    auto image = [&](uint32_t parity, uint32_t slot, int which) {
        return mb.images.as<float>() + mb.pixels * ((((size_t)parity * mb.bmax + slot) * 2) + which);
    };
    const VolumeProperties vp = whole_volume(op.getReconstructedVolume()->properties());
    std::vector<dev::FusedBp> pending;   // back projections of the previous block
    auto flush_bp_only = [&](size_t keep) {
        // emit BP-only passes until at most `keep` back projections are pending
        while (pending.size() > keep) {
            dev::FusedPass p{};
            p.zdir = 1;
            const size_t n = std::min(pending.size() - keep, (size_t)dev::kMaxFusedBp);
            for (size_t i = 0; i < n; ++i) p.bp[p.nbp++] = pending[i];
            pending.erase(pending.begin(), pending.begin() + n);
            mb.passes.push_back(p);
        }
    };
    const uint32_t nb = op.GetBlockCount();
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t bs = op.getBlockSize(b);
        const uint32_t parity = b & 1u;
        // long-object compensation uses the block's last geometry (compensateLongObject)
        const ParallelScannerGeometry& last = op.geometry(op.getPhysicalProjectionIndex(b, bs - 1));
        const float loc_c = mb.use_loc ? last.traversal_length_through(vp.bbox_min(), vp.bbox_max()) : 0.0f;
        std::vector<dev::FusedFp> neg, pos;
        for (uint32_t p = 0; p < bs; ++p) {
            const uint32_t phys = op.getPhysicalProjectionIndex(b, p);
            const ParallelScannerGeometry& g = op.geometry(phys);
            dev::FusedFp f{};
            g.source_base().store(f.src);
            g.horizontal_pitch().store(f.h);
            g.vertical_pitch().store(f.v);
            g.ray_direction().store(f.ray);
            f.measured = op.realProjection(phys);
            f.residual = image(parity, p, 0);
            f.length = image(parity, p, 1);
            f.loc_c = loc_c;
            f.use_loc = mb.use_loc ? 1 : 0;
            (f.ray[2] > 0.0f ? pos : neg).push_back(f);
        }
        // FP groups: rays moving towards -z (or staying in z) and towards +z
        std::vector<std::pair<int, std::vector<dev::FusedFp>>> groups;
        for (auto* grp : {&neg, &pos}) {
            const int zdir = grp == &pos ? 1 : -1;
            for (size_t i = 0; i < grp->size(); i += (size_t)max_fp)
                groups.push_back({zdir, std::vector<dev::FusedFp>(grp->begin() + i,
                                                                  grp->begin() + std::min(grp->size(), i + (size_t)max_fp))});
        }
        flush_bp_only(dev::kMaxFusedBp);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            dev::FusedPass p{};
            p.zdir = groups[gi].first;
            for (const auto& f : groups[gi].second) p.fp[p.nfp++] = f;
            if (gi == 0)
                for (const auto& e : pending) p.bp[p.nbp++] = e;
            mb.passes.push_back(p);
        }
        pending.clear();
        for (uint32_t p = 0; p < bs; ++p) {
            const ParallelScannerGeometry& g = op.geometry(op.getPhysicalProjectionIndex(b, p));
            dev::FusedBp e{};
            std::memcpy(e.m, g.projection_matrix().data(), sizeof(e.m));
            e.lambda = mb.lambda / (float)bs;
            e.residual = image(parity, p, 0);
            e.length = image(parity, p, 1);
            const std::vector<VoxelSample> smp = bp_sample_pattern(g, vp.voxel_size(), (uint32_t)mb.bp_samples);
            e.sample = make_float4(smp[0].dx, smp[0].dy, smp[0].dz, smp[0].length);
            e.nsamples = (int)smp.size();
            e.per_dir = mb.rowsets;
            for (size_t k = 0; k < smp.size(); ++k)
                e.samples[k] = make_float4(smp[k].dx, smp[k].dy, smp[k].dz, smp[k].length);
            pending.push_back(e);
        }
    }
    flush_bp_only(0);
}

class FusedEngineImpl final : public FusedPlaneEngine {
public:
    FusedEngineImpl(std::vector<MemberPlan> members, etn_voxel_type vt, bool vec, bool cap, bool exact, bool mask,
                    bool over, hipStream_t stream);
    void enqueueIteration(hipStream_t stream) override;
    void check(hipStream_t stream) override;
    std::vector<LaunchTime> timeIteration(hipStream_t stream) override;
    void autotune(hipStream_t stream) override;
    void setTile(size_t member, int planes, int band, hipStream_t stream) override;
    void tile(size_t member, int* planes, int* band) const override;
    void setThreads(int threads, hipStream_t stream) override;
    int threads() const override { return threads_; }
    size_t memberCount() const override { return members_.size(); }
    double residualNorm(size_t member, double* scratch, hipStream_t stream) override;

private:
    struct Launch {
        size_t work;            // first work item
        int groups;
        int rpt;
        size_t lds;
        bool steady;
    };
    struct Schedule {
        std::vector<Launch> launches;
        DeviceBuffer work, tables;
    };
    void upload(hipStream_t stream);
    void level_offsets();
    void build_schedule(const std::vector<size_t>& ids, Schedule& out, hipStream_t stream) const;
    void enqueue(const Schedule& s, size_t first, size_t last, hipStream_t stream, std::vector<hipEvent_t>* ev) const;
    bool threadsFit(int threads) const;
    void useThreads(int threads);
    void tuneTiles(hipStream_t stream);
    float timeLaunches(const Schedule& s, hipStream_t stream) const;

    std::vector<MemberPlan> members_;
    etn_voxel_type vt_;
    bool vec_, cap_, exact_;
    bool mask_;                         // every member has a mask (the launches use the mask kernels)
    bool over_;                         // every member oversamples (the launches use the oversampled kernels)
    int threads_ = dev::kMaxThreads;    // workgroup shape of every launch
    DeviceBuffer diag_, vols_, passes_;
    std::vector<size_t> pass_base_;     // each member's first pass in passes_
    Schedule sched_;
};

FusedEngineImpl::FusedEngineImpl(std::vector<MemberPlan> members, etn_voxel_type vt, bool vec, bool cap, bool exact,
                                 bool mask, bool over, hipStream_t stream)
    : members_(std::move(members)), vt_(vt), vec_(vec), cap_(cap), exact_(exact), mask_(mask), over_(over)
{
    diag_.allocate(sizeof(int));
    ETN_HIP(hipMemsetAsync(diag_.get(), 0, sizeof(int), stream));
    upload(stream);
}

// Passes and volume descriptors of every member under the current tiles, and
// the schedule of all members.
void FusedEngineImpl::upload(hipStream_t stream)
{
    std::vector<dev::FusedPass> all;
    std::vector<dev::FusedVolume> vols;
    pass_base_.clear();
    for (MemberPlan& m : members_) {
        build_passes(m);
        pass_base_.push_back(all.size());
        all.insert(all.end(), m.passes.begin(), m.passes.end());
        GpuVolume* vol = m.op->getReconstructedVolume();
        const VolumeProperties vp = whole_volume(vol->properties());
        dev::FusedVolume fv{};
        fv.data = vol->data();
        fv.slab = make_slab_params(vp, 0);
        fv.nx = m.nx;
        fv.ny = m.ny;
        fv.nz = m.nz;
        fv.nu = m.nu;
        fv.nv = m.nv;
        fv.planes = m.tile.planes;
        fv.band = m.tile.band;
        fv.rows_per_plane = m.rows_per_plane;
        fv.row_len = m.row_len;
        fv.pad_l = m.pad_l;
        fv.row_lo = m.rows.as<int>();
        fv.row_hi = m.rows.as<int>() + m.ny;
        fv.dead_rows = m.rows.as<int>() + 2 * m.ny;
        fv.ndead = m.ndead;
        fv.mask = m.mask;
        fv.fp_n = m.fp_n;
        fv.fp_shift = m.fp_shift;
        fv.rowsets = m.rowsets;
        fv.diag = diag_.as<int>();
        vols.push_back(fv);
    }
    ETN_HIP(hipStreamSynchronize(stream));   // no launch may still read the old tables
    passes_.allocate(sizeof(dev::FusedPass) * all.size());
    vols_.allocate(sizeof(dev::FusedVolume) * vols.size());
    ETN_HIP(hipMemcpyAsync(passes_.get(), all.data(), passes_.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(vols_.get(), vols.data(), vols_.bytes(), hipMemcpyHostToDevice, stream));
    level_offsets();
    std::vector<size_t> ids(members_.size());
    std::iota(ids.begin(), ids.end(), 0);
    build_schedule(ids, sched_, stream);
    ETN_HIP(hipStreamSynchronize(stream));
}

// Members with fewer passes than the longest start later or earlier so that
// the voxels a launch covers stay as even as possible across launches: each
// member, largest total first, takes the offset whose busiest launch is
// least loaded.
void FusedEngineImpl::level_offsets()
{
    size_t length = 0;
    for (const MemberPlan& m : members_) length = std::max(length, m.passes.size());
    std::vector<size_t> order(members_.size());
    std::iota(order.begin(), order.end(), 0);
    auto voxels = [&](size_t i) { return (double)members_[i].nx * members_[i].ny * members_[i].nz; };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return voxels(a) * members_[a].passes.size() > voxels(b) * members_[b].passes.size();
    });
    std::vector<double> load(length, 0.0);
    for (size_t i : order) {
        const size_t n = members_[i].passes.size();
        size_t best = 0;
        double best_peak = 0.0, best_sum = 0.0;
        for (size_t o = 0; o + n <= length; ++o) {
            double peak = 0.0, sum = 0.0;
            for (size_t k = o; k < o + n; ++k) {
                peak = std::max(peak, load[k]);
                sum += load[k];
            }
            if (o == 0 || peak < best_peak || (peak == best_peak && sum < best_sum)) {
                best = o;
                best_peak = peak;
                best_sum = sum;
            }
        }
        members_[i].offset = (int)best;
        for (size_t k = best; k < best + n; ++k) load[k] += voxels(i);
    }
}

// Launch k carries pass k - offset of every member in `ids` that has one,
// members with the most work per workgroup first.
void FusedEngineImpl::build_schedule(const std::vector<size_t>& ids, Schedule& out, hipStream_t stream) const
{
    size_t length = 0;
    for (size_t i : ids) length = std::max(length, (size_t)members_[i].offset + members_[i].passes.size());
    std::vector<size_t> order(ids);
    auto group_work = [&](size_t i) {
        const MemberPlan& m = members_[i];
        return (double)m.tile.planes * m.nx * m.nz;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return group_work(a) > group_work(b); });
    const dev::FusedPass* base = passes_.as<dev::FusedPass>();
    std::vector<dev::FusedWork> work;
    std::vector<const dev::FusedPass*> tables(length * members_.size(), nullptr);
    out.launches.clear();
    for (size_t k = 0; k < length; ++k) {
        Launch l{work.size(), 0, 1, 0, true};
        for (size_t i : order) {
            const MemberPlan& m = members_[i];
            if (k < (size_t)m.offset || k >= (size_t)m.offset + m.passes.size()) continue;
            const size_t pi = k - m.offset;
            const dev::FusedPass& p = m.passes[pi];
            tables[k * members_.size() + i] = base + pass_base_[i] + pi;
            for (int j = 0; j < m.ny; j += m.tile.planes) work.push_back({(int)i, j});
            l.rpt = std::max(l.rpt, select_rpt((p.nfp * m.tile.planes * m.rows_per_plane * m.nu) << m.fp_shift,
                                               threads_));
            l.lds = std::max(l.lds, sizeof(float) * (size_t)dev::fused_lds_floats(m.tile.planes, m.tile.band, m.nx,
                                                                                   m.row_floats, mask_, threads_,
                                                                                   m.rowsets));
            l.steady = l.steady && p.nfp > 0 && p.nbp > 0;
        }
        l.groups = (int)(work.size() - l.work);
        out.launches.push_back(l);
    }
    ETN_HIP(hipStreamSynchronize(stream));
    out.work.allocate(sizeof(dev::FusedWork) * std::max<size_t>(work.size(), 1));
    out.tables.allocate(sizeof(const dev::FusedPass*) * std::max<size_t>(tables.size(), 1));
    ETN_HIP(hipMemcpyAsync(out.work.get(), work.data(), sizeof(dev::FusedWork) * work.size(), hipMemcpyHostToDevice,
                           stream));
    ETN_HIP(hipMemcpyAsync(out.tables.get(), tables.data(), sizeof(const dev::FusedPass*) * tables.size(),
                           hipMemcpyHostToDevice, stream));
    ETN_HIP(hipStreamSynchronize(stream));
}

void FusedEngineImpl::enqueue(const Schedule& s, size_t first, size_t last, hipStream_t stream,
                              std::vector<hipEvent_t>* ev) const
{
    const dev::FusedWork* work = s.work.as<dev::FusedWork>();
    const dev::FusedPass* const* tables = s.tables.as<const dev::FusedPass*>();
    for (size_t k = first; k < last; ++k) {
        const Launch& l = s.launches[k];
        if (l.groups == 0) continue;
        const dev::FusedLaunch fl{threads_, l.rpt, l.groups, vols_.as<dev::FusedVolume>(), work + l.work,
                                  tables + k * members_.size(), l.lds};
        if (ev) ETN_HIP(hipEventRecord((*ev)[2 * k], stream));
        dev::launch_fused(vt_, vec_, cap_, exact_, mask_, over_, fl, stream);
        if (ev) ETN_HIP(hipEventRecord((*ev)[2 * k + 1], stream));
    }
}

void FusedEngineImpl::enqueueIteration(hipStream_t stream) { enqueue(sched_, 0, sched_.launches.size(), stream, nullptr); }

double FusedEngineImpl::residualNorm(size_t member, double* scratch, hipStream_t stream)
{
    if (member >= members_.size()) fail("member index out of range");
    const MemberPlan& m = members_[member];
    BlockIterativeReconstructionOperator& op = *m.op;
    const std::vector<uint32_t>& measured = op.measuredProjections();
    const size_t M = members_.size();
    const int max_fp = std::max(1, std::min(dev::kMaxFusedFp, threads_ * dev::kMaxRaysPerThread /
                                                                  ((m.tile.planes * m.rows_per_plane * m.nu)
                                                                   << m.fp_shift)));
    // virtual projection and ray length images for the projections of one chunk
    DeviceBuffer images(sizeof(float) * m.pixels * 2 * (size_t)max_fp);
    auto slot = [&](int k, int which) { return images.as<float>() + m.pixels * (2 * (size_t)k + which); };
    // chunks of max_fp measured projections: forward-only passes by z direction
    std::vector<dev::FusedPass> passes;
    std::vector<size_t> chunk_end;     // passes of chunk c end at chunk_end[c]
    for (size_t c0 = 0; c0 < measured.size(); c0 += (size_t)max_fp) {
        dev::FusedPass neg{}, pos{};
        neg.zdir = -1;
        pos.zdir = 1;
        for (size_t k = c0; k < std::min(measured.size(), c0 + (size_t)max_fp); ++k) {
            const ParallelScannerGeometry& g = op.geometry(measured[k]);
            dev::FusedFp f{};
            g.source_base().store(f.src);
            g.horizontal_pitch().store(f.h);
            g.vertical_pitch().store(f.v);
            g.ray_direction().store(f.ray);
            f.measured = op.realProjection(measured[k]);
            f.residual = slot((int)(k - c0), 0);
            f.length = slot((int)(k - c0), 1);
            f.raw = 1;
            dev::FusedPass& p = f.ray[2] > 0.0f ? pos : neg;
            p.fp[p.nfp++] = f;
        }
        if (neg.nfp) passes.push_back(neg);
        if (pos.nfp) passes.push_back(pos);
        chunk_end.push_back(passes.size());
    }
    std::vector<dev::FusedWork> work;
    for (int j = 0; j < m.ny; j += m.tile.planes) work.push_back({(int)member, j});
    DeviceBuffer dpasses(sizeof(dev::FusedPass) * std::max<size_t>(passes.size(), 1));
    DeviceBuffer dwork(sizeof(dev::FusedWork) * work.size());
    DeviceBuffer dtables(sizeof(const dev::FusedPass*) * std::max<size_t>(passes.size() * M, 1));
    std::vector<const dev::FusedPass*> tables(passes.size() * M, nullptr);
    for (size_t k = 0; k < passes.size(); ++k) tables[k * M + member] = dpasses.as<dev::FusedPass>() + k;
    ETN_HIP(hipMemcpyAsync(dpasses.get(), passes.data(), sizeof(dev::FusedPass) * passes.size(),
                           hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dwork.get(), work.data(), dwork.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dtables.get(), tables.data(), sizeof(const dev::FusedPass*) * tables.size(),
                           hipMemcpyHostToDevice, stream));
    double* partials = scratch;
    double* total = scratch + dev::kSsdBlocks;
    ETN_HIP(hipMemsetAsync(total, 0, sizeof(double), stream));
    const size_t lds = sizeof(float) * (size_t)dev::fused_lds_floats(m.tile.planes, m.tile.band, m.nx, m.row_floats,
                                                                      mask_, threads_, m.rowsets);
    size_t pass = 0;
    for (size_t c = 0; c < chunk_end.size(); ++c) {
        for (; pass < chunk_end[c]; ++pass) {
            const int rays = (passes[pass].nfp * m.tile.planes * m.rows_per_plane * m.nu) << m.fp_shift;
            const dev::FusedLaunch fl{threads_, select_rpt(rays, threads_), (int)work.size(),
                                      vols_.as<dev::FusedVolume>(), dwork.as<dev::FusedWork>(),
                                      dtables.as<const dev::FusedPass*>() + pass * M, lds};
            dev::launch_fused(vt_, vec_, cap_, exact_, mask_, over_, fl, stream);
        }
        // squared differences in projection order, as the unfused norm adds them
        const size_t c0 = c * (size_t)max_fp;
        for (size_t k = c0; k < std::min(measured.size(), c0 + (size_t)max_fp); ++k)
            dev::launch_sum_squared_difference(op.realProjection(measured[k]), slot((int)(k - c0), 0), m.pixels,
                                               partials, total, stream);
    }
    double host = 0.0;
    ETN_HIP(hipMemcpyAsync(&host, total, sizeof(double), hipMemcpyDeviceToHost, stream));
    ETN_HIP(hipStreamSynchronize(stream));
    return std::sqrt(host);
}

void FusedEngineImpl::check(hipStream_t stream)
{
    int bad = 0;
    ETN_HIP(hipMemcpyAsync(&bad, diag_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
    ETN_HIP(hipStreamSynchronize(stream));
    if (bad) throw Error(ETN_ERROR_INTERNAL, "fused plane kernel found rays outside their plane");
}

std::vector<FusedPlaneEngine::LaunchTime> FusedEngineImpl::timeIteration(hipStream_t stream)
{
    const size_t n = sched_.launches.size();
    std::vector<hipEvent_t> ev(2 * n);
    for (auto& e : ev) ETN_HIP(hipEventCreate(&e));
    enqueue(sched_, 0, n, stream, &ev);
    ETN_HIP(hipStreamSynchronize(stream));
    std::vector<LaunchTime> out;
    for (size_t k = 0; k < n; ++k) {
        if (sched_.launches[k].groups == 0) continue;
        float ms = 0.0f;
        ETN_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        out.push_back({ms, sched_.launches[k].steady});
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return out;
}

void FusedEngineImpl::setTile(size_t member, int planes, int band, hipStream_t stream)
{
    if (member >= members_.size()) fail("member index out of range");
    const Tile t{planes, band};
    if (!tile_fits(members_[member], t)) fail("tile does not fit this volume (planes must divide ny, LDS limit)");
    members_[member].tile = t;
    upload(stream);
}

void FusedEngineImpl::tile(size_t member, int* planes, int* band) const
{
    if (member >= members_.size()) fail("member index out of range");
    *planes = members_[member].tile.planes;
    *band = members_[member].tile.band;
}

// Every member keeps a fused form with `threads` per workgroup: its rays and
// a z row of one plane fit.
bool FusedEngineImpl::threadsFit(int threads) const
{
    if (threads != 256 && threads != 512) return false;
    for (const MemberPlan& m : members_)
        if (max_band(m, 1, threads) < 1) return false;
    return true;
}

// Switches every member to `threads`; tiles that no longer fit fall back to the default.
void FusedEngineImpl::useThreads(int threads)
{
    threads_ = threads;
    for (MemberPlan& m : members_) {
        m.threads = threads;
        if (!tile_fits(m, m.tile)) m.tile = default_tile(m);
    }
}

void FusedEngineImpl::setThreads(int threads, hipStream_t stream)
{
    if (!threadsFit(threads)) fail("workgroup size must be 256 or 512 threads and fit every member");
    useThreads(threads);
    upload(stream);
}

// Event time of launches 1 .. kTimed of a schedule (launch 0 runs untimed); -1
// when the schedule has fewer than two launches.
float FusedEngineImpl::timeLaunches(const Schedule& s, hipStream_t stream) const
{
    constexpr size_t kTimed = 4;
    const size_t n = s.launches.size();
    if (n < 2) return -1.0f;
    hipEvent_t a, b;
    ETN_HIP(hipEventCreate(&a));
    ETN_HIP(hipEventCreate(&b));
    enqueue(s, 0, 1, stream, nullptr);
    ETN_HIP(hipEventRecord(a, stream));
    enqueue(s, 1, std::min(n, 1 + kTimed), stream, nullptr);
    ETN_HIP(hipEventRecord(b, stream));
    ETN_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    ETN_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

// Members of one size share a tile; each size is timed on its own, under the
// current workgroup shape.
void FusedEngineImpl::tuneTiles(hipStream_t stream)
{
    // (the staged row length and row sets and the ray slots per pixel decide
    // which tiles fit as much as the sizes do)
    std::map<std::tuple<int, int, int, int, int, int, int, int, int>, std::vector<size_t>> classes;
    for (size_t i = 0; i < members_.size(); ++i) {
        const MemberPlan& m = members_[i];
        classes[{m.nx, m.ny, m.nz, m.nu, m.nv, m.rows_per_plane, m.row_floats, m.rowsets, m.fp_shift}].push_back(i);
    }
    for (auto& kv : classes) {
        const std::vector<size_t>& ids = kv.second;
        const std::vector<Tile> cands = candidate_tiles(members_[ids[0]]);
        if (cands.size() <= 1) continue;
        float best_ms = -1.0f;
        Tile best = members_[ids[0]].tile;
        for (const Tile& t : cands) {
            for (size_t i : ids) members_[i].tile = t;
            upload(stream);
            for (size_t i : ids) members_[i].offset = 0;
            Schedule s;
            build_schedule(ids, s, stream);
            const float ms = timeLaunches(s, stream);
            if (ms >= 0.0f && (best_ms < 0.0f || ms < best_ms)) {
                best_ms = ms;
                best = t;
            }
        }
        for (size_t i : ids) members_[i].tile = best;
    }
}

// For each workgroup shape that fits every member: tune the tiles, then time
// the engine's whole schedule; keep the fastest shape with its tiles.
void FusedEngineImpl::autotune(hipStream_t stream)
{
    std::vector<int> shapes;
    for (int t : {512, 256})
        if (threadsFit(t)) shapes.push_back(t);
    bool any = shapes.size() > 1;
    for (const MemberPlan& m : members_) any = any || candidate_tiles(m).size() > 1;
    if (!any) {
        upload(stream);
        return;
    }
    // save the volumes: the trial passes update them
    std::vector<size_t> bytes(members_.size());
    size_t total = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        bytes[i] = members_[i].op->getReconstructedVolume()->bytes();
        total += bytes[i];
    }
    DeviceBuffer saved(total);
    size_t at = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        ETN_HIP(hipMemcpyAsync(saved.as<char>() + at, members_[i].op->getReconstructedVolume()->data(), bytes[i],
                               hipMemcpyDeviceToDevice, stream));
        at += bytes[i];
    }
    int best_threads = threads_;
    std::vector<Tile> best_tiles;
    float best_ms = -1.0f;
    for (int t : shapes) {
        useThreads(t);
        for (MemberPlan& m : members_) m.tile = default_tile(m);
        tuneTiles(stream);
        upload(stream);
        const float ms = shapes.size() > 1 ? timeLaunches(sched_, stream) : 0.0f;
        if (best_tiles.empty() || (ms >= 0.0f && ms < best_ms)) {
            best_ms = ms;
            best_threads = t;
            best_tiles.clear();
            for (const MemberPlan& m : members_) best_tiles.push_back(m.tile);
        }
    }
    useThreads(best_threads);
    for (size_t i = 0; i < members_.size(); ++i) members_[i].tile = best_tiles[i];
    at = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        ETN_HIP(hipMemcpyAsync(members_[i].op->getReconstructedVolume()->data(), saved.as<char>() + at, bytes[i],
                               hipMemcpyDeviceToDevice, stream));
        at += bytes[i];
    }
    upload(stream);
    check(stream);
}

}  // namespace

namespace {

class FusedBackProjectorImpl final : public FusedBackProjector {
public:
    FusedBackProjectorImpl(GpuVolume* volume, const std::vector<ParallelScannerGeometry>& geometries, MemberPlan plan)
        : volume_(volume), geometries_(geometries), m_(std::move(plan))
    {
        diag_.allocate(sizeof(int));
        ETN_HIP(hipMemset(diag_.get(), 0, sizeof(int)));
    }
    void backProject(const std::vector<uint32_t>& indices, const std::vector<const float*>& residual,
                     const std::vector<const float*>& length, float lambda, hipStream_t stream,
                     std::vector<float>* launch_ms) override;

    GpuVolume* volume_;
    std::vector<ParallelScannerGeometry> geometries_;
    MemberPlan m_;
    DeviceBuffer diag_;
};

void FusedBackProjectorImpl::backProject(const std::vector<uint32_t>& indices,
                                         const std::vector<const float*>& residual,
                                         const std::vector<const float*>& length, float lambda, hipStream_t stream,
                                         std::vector<float>* launch_ms)
{
    if (residual.size() != indices.size() || length.size() != indices.size())
        fail("fused back projector: one residual and one length image per projection");
    const VolumeProperties vp = whole_volume(volume_->properties());
    std::vector<dev::FusedPass> passes;
    for (size_t k = 0; k < indices.size(); ++k) {
        if (indices[k] >= geometries_.size()) fail("fused back projector: projection index out of range");
        if (k % dev::kMaxFusedBp == 0) {
            passes.push_back(dev::FusedPass{});
            passes.back().zdir = 1;
        }
        const ParallelScannerGeometry& g = geometries_[indices[k]];
        dev::FusedBp e{};
        std::memcpy(e.m, g.projection_matrix().data(), sizeof(e.m));
        e.lambda = lambda;
        e.residual = residual[k];
        e.length = length[k];
        const std::vector<VoxelSample> smp = bp_sample_pattern(g, vp.voxel_size(), (uint32_t)m_.bp_samples);
        e.sample = make_float4(smp[0].dx, smp[0].dy, smp[0].dz, smp[0].length);
        e.nsamples = (int)smp.size();
        e.per_dir = m_.rowsets;
        for (size_t t = 0; t < smp.size(); ++t) e.samples[t] = make_float4(smp[t].dx, smp[t].dy, smp[t].dz, smp[t].length);
        dev::FusedPass& p = passes.back();
        p.bp[p.nbp++] = e;
    }
    if (passes.empty()) return;
    dev::FusedVolume fv{};
    fv.data = volume_->data();
    fv.slab = make_slab_params(vp, 0);
    fv.nx = m_.nx;
    fv.ny = m_.ny;
    fv.nz = m_.nz;
    fv.nu = m_.nu;
    fv.nv = m_.nv;
    fv.planes = m_.tile.planes;
    fv.band = m_.tile.band;
    fv.rows_per_plane = m_.rows_per_plane;
    fv.row_len = m_.row_len;
    fv.pad_l = m_.pad_l;
    fv.fp_n = 1;
    fv.rowsets = m_.rowsets;
    fv.diag = diag_.as<int>();
    std::vector<dev::FusedWork> work;
    for (int j = 0; j < m_.ny; j += m_.tile.planes) work.push_back({0, j});
    DeviceBuffer dvol(sizeof(dev::FusedVolume)), dpasses(sizeof(dev::FusedPass) * passes.size());
    DeviceBuffer dwork(sizeof(dev::FusedWork) * work.size()), dtables(sizeof(const dev::FusedPass*) * passes.size());
    std::vector<const dev::FusedPass*> tables(passes.size());
    for (size_t k = 0; k < passes.size(); ++k) tables[k] = dpasses.as<dev::FusedPass>() + k;
    ETN_HIP(hipMemcpyAsync(dvol.get(), &fv, sizeof(fv), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dpasses.get(), passes.data(), dpasses.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dwork.get(), work.data(), dwork.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dtables.get(), tables.data(), dtables.bytes(), hipMemcpyHostToDevice, stream));
    const size_t lds = sizeof(float) * (size_t)dev::fused_lds_floats(m_.tile.planes, m_.tile.band, m_.nx,
                                                                      m_.row_floats, false, m_.threads, m_.rowsets);
    std::vector<hipEvent_t> ev(launch_ms ? 2 * passes.size() : 0);
    for (auto& e : ev) ETN_HIP(hipEventCreate(&e));
    for (size_t k = 0; k < passes.size(); ++k) {
        const dev::FusedLaunch fl{m_.threads, 1, (int)work.size(), dvol.as<dev::FusedVolume>(),
                                  dwork.as<dev::FusedWork>(), dtables.as<const dev::FusedPass*>() + k, lds};
        if (launch_ms) ETN_HIP(hipEventRecord(ev[2 * k], stream));
        dev::launch_fused(volume_->voxel_type(), true, false, false, false, m_.oversampled(), fl, stream);
        if (launch_ms) ETN_HIP(hipEventRecord(ev[2 * k + 1], stream));
    }
    ETN_HIP(hipStreamSynchronize(stream));   // the tables above are freed on return
    for (size_t k = 0; launch_ms && k < passes.size(); ++k) {
        float ms = 0.0f;
        ETN_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        if (passes[k].nbp == dev::kMaxFusedBp) launch_ms->push_back(ms);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
}

}  // namespace

std::unique_ptr<FusedBackProjector> FusedBackProjector::create(GpuVolume* volume,
                                                               const std::vector<ParallelScannerGeometry>& geometries,
                                                               uint32_t nu, uint32_t nv, uint32_t bp_samples,
                                                               hipStream_t)
{
    const VolumeProperties vp = whole_volume(volume->properties());
    MemberPlan m;
    m.nx = (int)vp.nx();
    m.ny = (int)vp.ny();
    m.nz = (int)vp.nz();
    m.nu = (int)nu;
    m.nv = (int)nv;
    if (m.nx % 4 || geometries.empty()) return nullptr;
    const int bp_s = (int)(bp_samples ? bp_samples : 1);
    const int pd = (int)std::lround(std::sqrt((double)bp_s));
    if (bp_s > dev::kMaxFusedBpSamples || pd * pd != bp_s) return nullptr;
    std::vector<const ParallelScannerGeometry*> geoms;
    for (const ParallelScannerGeometry& g : geometries) {
        // the v coordinate of a voxel depends on its y only, w == 1
        if (g.ray_direction().y != 0.0f || g.horizontal_pitch().y != 0.0f) return nullptr;
        const Matrix4x4 pm = g.projection_matrix();
        const float* mv = pm.data();
        if (mv[1] != 0.0f || mv[9] != 0.0f || mv[3] != 0.0f || mv[7] != 0.0f || mv[11] != 0.0f || mv[15] != 1.0f)
            return nullptr;
        geoms.push_back(&g);
    }
    if (!fast_row_span(vp, geoms, m.nu, bp_s, m.pad_l, m.row_len)) return nullptr;
    m.pixels = (size_t)nu * nv;
    m.rows_per_plane = 1;
    m.row_floats = m.row_len * 4;
    m.threads = dev::kMaxThreads;
    m.bp_samples = bp_s;
    m.rowsets = pd;
    if (dev::fused_lds_floats(1, 1, m.nx, m.row_floats, false, m.threads, m.rowsets) >
        dev::shape_lds_floats(m.threads) || max_band(m, 1) < 1)
        return nullptr;
    m.tile = default_tile(m);
    return std::make_unique<FusedBackProjectorImpl>(volume, geometries, std::move(m));
}

std::unique_ptr<FusedPlaneEngine> FusedPlaneEngine::create(const std::vector<Member>& members, bool exact,
                                                           hipStream_t stream, std::vector<bool>* taken)
{
    taken->assign(members.size(), false);
    std::vector<MemberPlan> plans;
    etn_voxel_type vt = ETN_VOXEL_FLOAT32;
    bool vec = true, cap = false, mask = false, over = false;
    for (size_t i = 0; i < members.size(); ++i) {
        const etn_reconstruction_desc& d = *members[i].desc;
        const bool d_over = (d.fp_samples ? d.fp_samples : 1) != 1 || (d.bp_samples ? d.bp_samples : 1) != 1;
        if (!plans.empty() && (d.voxel_type != vt || (d.mask != nullptr) != mask || d_over != over ||
                               (members[i].op->getReconstructedVolume()->properties().nx() % 4 == 0) != vec))
            continue;
        MemberPlan m;
        if (!plan_member(*members[i].op, d, exact, m)) continue;
        if (plans.empty()) {
            vt = d.voxel_type;
            vec = m.nx % 4 == 0;
            mask = m.mask != nullptr;
            over = m.oversampled();
        }
        cap = cap || m.nx + m.nz > 2049;
        plans.push_back(std::move(m));
        (*taken)[i] = true;
    }
    if (plans.empty()) return nullptr;
    return std::make_unique<FusedEngineImpl>(std::move(plans), vt, vec, cap, exact, mask, over, stream);
}

}  // namespace etn
